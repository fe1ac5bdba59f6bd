// Host engine of the MI355X TAN convolution library: contexts (device +
// process/general streams), convolution objects (device-resident IR slots,
// FDL and tail partial sums; the IR-switch state machine of the reference's
// TANConvolutionImpl::Process) and the C ABI declared in tan_mi355x.h.
#include "tan_internal.h"
#include "../../include/tan_mi355x.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

using namespace tanx;

namespace {

int32_t map_hip(hipError_t e)
{
    if (e == hipSuccess) return TANX_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return TANX_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return TANX_NO_DEVICE;
    if (e == hipErrorInvalidValue) return TANX_INVALID_ARG;
    return TANX_UNEXPECTED;
}

#define TANX_CHECK(expr)                                          \
    do {                                                          \
        hipError_t _e = (expr);                                   \
        if (_e != hipSuccess) {                                   \
            std::fprintf(stderr, "tan_mi355x: %s failed: %s (%s:%d)\n", #expr, \
                         hipGetErrorString(_e), __FILE__, __LINE__); \
            return map_hip(_e);                                   \
        }                                                         \
    } while (0)

// Makes `dev` current for the scope and restores the caller's device.
struct DeviceGuard {
    int prev = -1;
    bool ok = false;
    explicit DeviceGuard(int dev)
    {
        // an error left by an earlier call (one the library handled, such as
        // a refused allocation, or the caller's own) must not be taken for
        // the result of this call's launches
        (void)hipGetLastError();
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        ok = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

inline bool is_pow2(uint32_t v) { return v && !(v & (v - 1)); }

}  // namespace

static constexpr int kTimerRing = 1024;

// Per-kernel durations measured with HIP events recorded on the process
// stream around each launch; drained into running sums when the ring fills
// or when read.
struct KernelTimer {
    std::vector<hipEvent_t> start, stop;
    int head = 0;     // slot the next launch records into
    int pending = 0;  // slots recorded and not read yet; the oldest is head - pending
    double ms = 0.0;
    uint64_t count = 0;
    bool init()
    {
        start.resize(kTimerRing);
        stop.resize(kTimerRing);
        for (int i = 0; i < kTimerRing; i++)
            if (hipEventCreate(&start[i]) != hipSuccess || hipEventCreate(&stop[i]) != hipSuccess) return false;
        return true;
    }
    void destroy()
    {
        for (auto e : start) if (e) (void)hipEventDestroy(e);
        for (auto e : stop) if (e) (void)hipEventDestroy(e);
        start.clear();
        stop.clear();
    }
    // reads the oldest recorded slot: waiting for it, or only if it is done
    // (hipErrorNotReady otherwise)
    hipError_t read_oldest(bool wait)
    {
        const int i = (head - pending + kTimerRing) % kTimerRing;
        hipError_t e = wait ? hipEventSynchronize(stop[i]) : hipEventQuery(stop[i]);
        if (e == hipErrorNotReady) {
            (void)hipGetLastError();  // not an error of the launches that follow
            return e;
        }
        if (e != hipSuccess) return e;
        float t = 0.f;
        e = hipEventElapsedTime(&t, start[i], stop[i]);
        if (e != hipSuccess) return e;
        ms += t;
        count++;
        pending--;
        return hipSuccess;
    }
    // events for the next launch.  A few completed slots are read first (in a
    // stream of launches the oldest are long done: a few queries per launch),
    // so no launch has to read a whole ring; only a full ring waits.
    hipError_t acquire(hipEvent_t *e0, hipEvent_t *e1)
    {
        for (int k = 0; k < 4 && pending > 0; k++) {
            const hipError_t e = read_oldest(false);
            if (e == hipErrorNotReady) break;
            if (e != hipSuccess) return e;
        }
        if (pending == kTimerRing) {
            const hipError_t e = read_oldest(true);
            if (e != hipSuccess) return e;
        }
        *e0 = start[head];
        *e1 = stop[head];
        return hipSuccess;
    }
    void commit()
    {
        head = (head + 1) % kTimerRing;
        pending++;
    }
    hipError_t drain()
    {
        while (pending > 0) {
            const hipError_t e = read_oldest(true);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
};

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct tanx_ctx {
    // the caller's handle and every object made on the context hold a
    // reference (TANContext is reference counted in the reference's AMF
    // object model): tanx_ctx_destroy drops the caller's, and the context
    // goes with the last one
    std::atomic<int> refs{1};
    int device = 0;
    // pass rows (FDL and two IR slots) of every live object on the context:
    // multi-block passes read them with ordinary loads only while the total
    // fits the cached bound, so objects do not evict one another
    std::atomic<uint64_t> pass_rows{0};
    // group launches: descriptors + block offsets, pinned ring and device copy
    std::mutex group_mtx;
    char *g_host = nullptr;     // [kGroupRing][g_cap]
    char *g_dev = nullptr;      // [g_cap]
    size_t g_cap = 0;
    hipEvent_t g_ev[4] = {};
    int g_ring = 0;
    // group-launch timing (tanx_ctx_enable_group_timing): launch-stamped
    // events around the tick and frame kernels, and the algorithmic bytes of
    // the roles they ran
    bool gtiming = false;
    KernelTimer gt_tick, gt_frame;
    uint64_t g_tick_bytes = 0, g_frame_bytes = 0, g_objects = 0;
    // group calls as work-queue launches (tanx_ctx_set_group_queue): one
    // counter per launch of a call, zeroed on the process stream before it
    bool gqueue = false;
    int *q_next = nullptr;
    hipStream_t proc = nullptr;  // convolution queue equivalent
    hipStream_t upd = nullptr;   // general queue equivalent (IR updates, FFT)
    bool own_proc = true, own_upd = true;  // created here (else the caller's)
    float2 *tw = nullptr;        // twiddle tables, see tw_table()
    std::mutex fft_mtx;
    float *fft_dev = nullptr;
    size_t fft_dev_floats = 0;
    float2 *fft_scratch = nullptr;  // four-step transforms: 2 x M complex per channel
    size_t fft_scratch_elems = 0;
    float *fft_host = nullptr;
    size_t fft_host_floats = 0;
    // matrix-core DFT tables per size: [n+2][n] bf16 high and low parts
    std::mutex dft_mtx;
    int dft_n = 0;
    uint16_t *dft_hi = nullptr, *dft_lo = nullptr;
    // output chain: clip flag and host staging (converter, host mixer)
    std::mutex io_mtx;
    int *io_flag = nullptr;       // device
    int *io_flag_host = nullptr;  // pinned
    char *io_dev = nullptr;
    char *io_host = nullptr;
    size_t io_cap = 0;
    // room responses: fixed-point accumulator (kept zero between calls by
    // the fill kernel), per-ear arguments (pinned staging + device copy) and
    // the host path's output buffers
    std::mutex room_mtx;
    unsigned *room_acc = nullptr;
    size_t room_acc_ints = 0;
    RoomEar *room_ears_dev = nullptr, *room_ears_host = nullptr;
    size_t room_ears_cap = 0;
    hipEvent_t room_ev = nullptr;  // staging copied to the device
    bool room_ev_live = false;
    float *room_out_dev = nullptr, *room_out_host = nullptr;
    size_t room_out_cap = 0;
};

static int32_t ctx_upload_twiddles(tanx_ctx *ctx)
{
    const size_t total = (size_t(1) << (kMaxTwLog2 + 1)) - 1;
    std::vector<float2> h(total);
    for (int l = 0; l <= kMaxTwLog2; l++) {
        const size_t n = size_t(1) << l, off = n - 1;
        for (size_t t = 0; t < n; t++) {
            // exp(-2 pi i t/n) in FP64, rounded once to FP32
            const double ang = -2.0 * M_PI * (double)t / (double)n;
            h[off + t] = make_float2((float)std::cos(ang), (float)std::sin(ang));
        }
    }
    TANX_CHECK(hipMalloc(&ctx->tw, total * sizeof(float2)));
    // on the process stream and waited for: a plain hipMemcpy from pageable
    // memory may return before its DMA lands, unordered with the context's
    // non-blocking streams
    TANX_CHECK(hipMemcpyAsync(ctx->tw, h.data(), total * sizeof(float2), hipMemcpyHostToDevice, ctx->proc));
    TANX_CHECK(hipStreamSynchronize(ctx->proc));
    return TANX_OK;
}

// A context on `device` whose process and general streams are the caller's
// (conv, general) or, where NULL, streams of its own.
static int32_t ctx_create(uint64_t version, int32_t device, hipStream_t conv, hipStream_t general, tanx_ctx **out)
{
    if (!out) return TANX_INVALID_POINTER;
    *out = nullptr;
    if ((version >> 48) != 1) return TANX_TAN_UNSUPPORTED_VERSION;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return TANX_NO_DEVICE;
    if (device < 0 || device >= ndev) return TANX_INVALID_ARG;
    DeviceGuard g(device);
    if (!g.ok) return TANX_NO_DEVICE;
    for (hipStream_t st : {conv, general}) {
        hipDevice_t d = -1;
        if (st && (hipStreamGetDevice(st, &d) != hipSuccess || d != device)) return TANX_INVALID_ARG;
    }
    tanx_ctx *ctx = new (std::nothrow) tanx_ctx();
    if (!ctx) return TANX_OUT_OF_MEMORY;
    ctx->device = device;
    ctx->proc = conv;
    ctx->upd = general;
    ctx->own_proc = conv == nullptr;
    ctx->own_upd = general == nullptr;
    int32_t rc = TANX_OK;
    // default-priority streams only
    if ((ctx->own_proc && hipStreamCreateWithFlags(&ctx->proc, hipStreamNonBlocking) != hipSuccess) ||
        (ctx->own_upd && hipStreamCreateWithFlags(&ctx->upd, hipStreamNonBlocking) != hipSuccess)) {
        rc = TANX_UNEXPECTED;
    }
    if (rc == TANX_OK) rc = ctx_upload_twiddles(ctx);
    // the device address of the allocation-free group entry point's scratch
    // array on this device (the device is current here)
    if (rc == TANX_OK) rc = map_hip(noalloc_scratch_init(device));
    if (rc != TANX_OK) {
        tanx_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return TANX_OK;
}

extern "C" int32_t tanx_ctx_create(uint64_t version, int32_t device, tanx_ctx **out)
{
    return ctx_create(version, device, nullptr, nullptr, out);
}

extern "C" int32_t tanx_ctx_create_on_streams(uint64_t version, int32_t device, void *conv_stream,
                                              void *general_stream, tanx_ctx **out)
{
    return ctx_create(version, device, static_cast<hipStream_t>(conv_stream),
                      static_cast<hipStream_t>(general_stream), out);
}

static void ctx_free(tanx_ctx *ctx)
{
    {
        DeviceGuard g(ctx->device);
        // the caller's streams are only drained, the context's own also destroyed
        if (ctx->proc) {
            (void)hipStreamSynchronize(ctx->proc);
            if (ctx->own_proc) (void)hipStreamDestroy(ctx->proc);
        }
        if (ctx->upd) {
            (void)hipStreamSynchronize(ctx->upd);
            if (ctx->own_upd) (void)hipStreamDestroy(ctx->upd);
        }
        if (ctx->tw) (void)hipFree(ctx->tw);
        if (ctx->q_next) (void)hipFree(ctx->q_next);
        if (ctx->fft_dev) (void)hipFree(ctx->fft_dev);
        if (ctx->fft_scratch) (void)hipFree(ctx->fft_scratch);
        if (ctx->fft_host) (void)hipHostFree(ctx->fft_host);
        if (ctx->g_host) (void)hipHostFree(ctx->g_host);
        if (ctx->g_dev) (void)hipFree(ctx->g_dev);
        for (auto &e : ctx->g_ev) if (e) (void)hipEventDestroy(e);
        ctx->gt_tick.destroy();
        ctx->gt_frame.destroy();
        if (ctx->dft_hi) (void)hipFree(ctx->dft_hi);
        if (ctx->dft_lo) (void)hipFree(ctx->dft_lo);
        if (ctx->io_flag) (void)hipFree(ctx->io_flag);
        if (ctx->io_flag_host) (void)hipHostFree(ctx->io_flag_host);
        if (ctx->io_dev) (void)hipFree(ctx->io_dev);
        if (ctx->io_host) (void)hipHostFree(ctx->io_host);
        if (ctx->room_acc) (void)hipFree(ctx->room_acc);
        if (ctx->room_ears_dev) (void)hipFree(ctx->room_ears_dev);
        if (ctx->room_ears_host) (void)hipHostFree(ctx->room_ears_host);
        if (ctx->room_ev) (void)hipEventDestroy(ctx->room_ev);
        if (ctx->room_out_dev) (void)hipFree(ctx->room_out_dev);
        if (ctx->room_out_host) (void)hipHostFree(ctx->room_out_host);
    }
    delete ctx;
}

static void ctx_retain(tanx_ctx *ctx) { ctx->refs.fetch_add(1, std::memory_order_relaxed); }
static void ctx_release(tanx_ctx *ctx)
{
    if (ctx->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) ctx_free(ctx);
}

extern "C" int32_t tanx_ctx_destroy(tanx_ctx *ctx)
{
    if (!ctx) return TANX_INVALID_POINTER;
    ctx_release(ctx);
    return TANX_OK;
}

extern "C" void *tanx_ctx_process_stream(tanx_ctx *ctx) { return ctx ? (void *)ctx->proc : nullptr; }
extern "C" void *tanx_ctx_update_stream(tanx_ctx *ctx) { return ctx ? (void *)ctx->upd : nullptr; }

extern "C" int32_t tanx_ctx_device(const tanx_ctx *ctx, int32_t *device)
{
    if (!ctx || !device) return TANX_INVALID_POINTER;
    *device = ctx->device;
    return TANX_OK;
}

// ---------------------------------------------------------------------------
// convolution object
// ---------------------------------------------------------------------------
static constexpr int kMaxChunks = 64;
static constexpr int kRunRing = 8;
static constexpr int kGraphBlocks = 64;
static constexpr int kDescRing = 4;


// While the pass rows of all live objects on a context total at most this
// (half the 256 MB Infinity Cache of an MI355X), k_window's multi-block
// passes read them with ordinary loads, which keep them resident between
// launches (about 8% more throughput at 64 channels x 65,536 taps); beyond
// it, and in every other launch, rows are streamed with nontemporal loads.
constexpr size_t kCachedRowsMax = (size_t)128 << 20;

struct tanx_conv {
    tanx_ctx *ctx = nullptr;
    std::mutex mtx;
    uint32_t method = 0;
    bool fmode = false;
    int L = 0, B = 0, C = 0, len = 0, P = 0;
    int nchunk = 1, pc = 1, maxchunk = 1;
    int tail_nchunk = 1;  // chunking of the partial sums the next head reads
    int tread = 0;        // partial-sum buffer the next head reads
    bool tail_due = false;  // a head ran since the last tail
    // Multi-block tail passes (kp > 1): a launch running the heads of blocks
    // [w, e) also runs, when block e has no partial sums yet (or a switch
    // armed at block e-1 made the old filter's stale), a pass computing the
    // sums of blocks e .. e+kp-1 from FDL rows before w.  The sums are kept
    // per block in a ring of 2*kp slots (block b: slot b mod 2kp), per
    // filter f (0: the block's filter, 1: the crossfade's second filter);
    // porg is the first head block of the launch that computed them (block
    // b's head adds partitions 0..b-porg itself), pchunk their chunk count,
    // covered the last block with sums.
    int kp = 1;
    int kphase = 0;  // passes end on blocks b with (b + kphase) % kp == 0
    // bytes of the rows a pass streams (FDL, both filters of a crossfade,
    // frame rows), counted in the context's pass_rows while the object lives
    uint64_t rows_bytes = 0;
    long long covered = 0;
    long long porg[2 * kPassMax][2] = {};
    int pchunk[2 * kPassMax][2] = {};
    bool fused = true;    // device path: head and next tail in one launch
    bool fixed_chunks = false;

    // device memory (one allocation carved into pieces)
    void *dmem = nullptr;
    size_t dbytes = 0;
    char *H = nullptr;       // [3][C][P][B] IR spectra of irtype
    float *S = nullptr;      // [3][C][P] row scales (reduced-precision spectra)
    int irtype = IR_F32;
    float2 *F = nullptr;     // [C][P][B]
    float2 *T = nullptr;     // tail partial sums [2 parity][2 filter][C][maxchunk][B]; with
                             // multi-block passes [2kp slots][2 filter][C][maxchunk][B]
    float *xprev = nullptr;  // [C][B]
    float *din = nullptr;    // [C][B]
    float *dout = nullptr;   // [C][B]
    float *dir = nullptr;    // [C][len] host-IR staging, allocated on first host update
    int *dchan = nullptr;    // [C]
    uint8_t *drun = nullptr; // [C]

    // pinned host staging
    float *hin = nullptr, *hout = nullptr, *hir = nullptr;
    int *hchan = nullptr;
    uint8_t *hrun = nullptr;  // [kRunRing][C]
    hipEvent_t run_ev[kRunRing] = {};
    int run_ring = 0;

    // IR updates run on the process stream, in order with every block
    // enqueued before and after them, so no event orders the two.  ev_upd
    // marks an update only when it used the pinned staging (host IRs, a
    // channel subset) or the caller blocks until it is ready.
    hipEvent_t ev_upd = nullptr;
    bool stage_busy = false;  // the last update's staging copies may still be in flight
    hipEvent_t ev_out = nullptr;   // host-path output copied

    // IR switch state (ConvolutionImpl.cpp:115-117, 869-927)
    int cur = 1, prev = 0, upd = 2;
    int nslots = 3;  // 2: the update slot is the previous filter's once a crossfade has run
    bool pending = false, ready = false, delayed = false, crossfading = false, last_xfade = false;
    int pos = 0;  // FDL ring slot of the next block
    std::vector<uint8_t> staged;  // channel staged since the last hand-off

    // channel flags (ConvolutionImpl.cpp:2868-2907)
    std::vector<uint8_t> available, flushed, run_now, run_dev;
    std::vector<int> tail_left;

    uint64_t ticks = 0, xfades = 0;

    bool timing = false;
    KernelTimer t_head, t_tail;

    // graph replay of device-memory blocks
    // opt-in: on this runtime a 64-node graph launch costs more host time
    // than 64 direct launches (about 22 us against 3 us per block)
    bool use_graph = false;
    struct GraphEntry {
        int k;
        unsigned blocks;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    TickArgs *d_desc = nullptr;   // [kGraphBlocks] launch descriptors (device)
    TickArgs *h_desc = nullptr;   // [kDescRing][kGraphBlocks] pinned staging
    hipEvent_t desc_ev[kDescRing] = {};
    int desc_ring = 0;
    uint64_t tail_alg_bytes = 0, head_alg_bytes = 0;
    uint64_t timed_blocks = 0;  // blocks whose heads the timed launches ran

    // time-domain objects: IR slots [3][C][len] and history [C][len] floats
    bool td = false;
    int tdpos = 0;  // history position of the next input sample
    float *td_slot(int s) const { return reinterpret_cast<float *>(H) + (size_t)s * C * len; }
    float *td_hist() const { return reinterpret_cast<float *>(F); }
    int *d_nz = nullptr;  // [C][2] ProcessDirect non-zero ranges (device)

    // non-uniform objects: frames of M blocks (N = M * B samples); the
    // uniform fields above then describe the first frame of taps (P = M)
    bool nu = false;
    int M = 1, N = 0, PB = 0, RB = 0;
    int nchB = 0, pcB = 1;
    char *HBm = nullptr;     // [3][C][PB][N] frame-partition spectra of irtype
    float *SBm = nullptr;    // [3][C][PB] row scales
    float2 *FB = nullptr;    // [C][RB][N] frame spectra ring
    float *XR = nullptr;     // [C][2M][B] input ring
    float2 *AC = nullptr;    // [2 parity][2 filter][C][nchB][N] frame partial sums
    float *ZB = nullptr;     // [2 parity][2 filter][C][N] frame outputs
    uint64_t switch_at = 0;  // block index of the armed crossfade
    int nu_shift = 0;        // frame clock = block index + nu_shift (frame phase option)
    long long zsel_frame = -1;  // frame whose current-filter output is in ZB[.][1]

    size_t brow_bytes() const { return (size_t)N * ir_elem_bytes(irtype); }
    size_t bslot_bytes() const { return (size_t)C * PB * brow_bytes(); }
    void *bslot_ptr(int s) const { return HBm + (size_t)s * bslot_bytes(); }
    float *bscale_ptr(int s) const { return SBm + (size_t)s * C * PB; }
    float2 *acc_ptr(int par, int f) const { return AC + ((size_t)par * 2 + f) * C * nchB * N; }
    float *zb_ptr(int par, int f) const { return ZB + ((size_t)par * 2 + f) * C * N; }

    size_t row_bytes() const { return (size_t)B * ir_elem_bytes(irtype); }
    size_t slot_bytes() const { return (size_t)C * P * row_bytes(); }
    void *slot_ptr(int s) const { return H + (size_t)s * slot_bytes(); }
    float *scale_ptr(int s) const { return S + (size_t)s * C * P; }
    // partial sums of buffer `par` for filter f (0 = current, 1 = next)
    float2 *tail_ptr(int par, int f) const { return T + ((size_t)par * 2 + f) * C * maxchunk * B; }
    // multi-block passes: ring slot of block b's partial sums
    int pring(long long b) const { return (int)(b % (2 * kp)); }
    float2 *tail_slot(long long b, int f) const { return T + ((size_t)pring(b) * 2 + f) * C * maxchunk * B; }
};

static void conv_free(tanx_conv *cv)
{
    if (!cv) return;
    if (cv->ctx) {
        cv->ctx->pass_rows.fetch_sub(cv->rows_bytes, std::memory_order_relaxed);
        DeviceGuard g(cv->ctx->device);
        (void)hipStreamSynchronize(cv->ctx->proc);
        (void)hipStreamSynchronize(cv->ctx->upd);
        if (cv->dmem) (void)hipFree(cv->dmem);
        if (cv->dir) (void)hipFree(cv->dir);
        if (cv->hin) (void)hipHostFree(cv->hin);
        if (cv->hout) (void)hipHostFree(cv->hout);
        if (cv->hir) (void)hipHostFree(cv->hir);
        if (cv->hchan) (void)hipHostFree(cv->hchan);
        if (cv->hrun) (void)hipHostFree(cv->hrun);
        for (auto &e : cv->run_ev) if (e) (void)hipEventDestroy(e);
        if (cv->ev_upd) (void)hipEventDestroy(cv->ev_upd);
        if (cv->ev_out) (void)hipEventDestroy(cv->ev_out);
        cv->t_head.destroy();
        cv->t_tail.destroy();
        for (auto &g : cv->graphs) (void)hipGraphExecDestroy(g.exec);
        if (cv->d_desc) (void)hipFree(cv->d_desc);
        if (cv->d_nz) (void)hipFree(cv->d_nz);
        if (cv->h_desc) (void)hipHostFree(cv->h_desc);
        for (auto &e : cv->desc_ev) if (e) (void)hipEventDestroy(e);
    }
    tanx_ctx *ctx = cv->ctx;
    delete cv;
    if (ctx) ctx_release(ctx);
}

static int auto_chunks(int C, int P)
{
    // enough workgroups to cover 256 CUs several times at small channel
    // counts, while every chunk streams at least 8 partitions (a chunk of
    // fewer rows is mostly launch and partial-sum overhead)
    int want = (1024 + C - 1) / C;
    want = std::max(1, std::min(want, 16));
    const int rows = std::max(1, (P - 2 + 7) / 8);
    return std::min(want, rows);
}

static void set_chunks(tanx_conv *cv, int n)
{
    const int parts = std::max(1, cv->P - 2);  // the tail covers partitions 2..P-1
    n = std::max(1, std::min(n, std::min(cv->maxchunk, parts)));
    cv->pc = (parts + n - 1) / n;
    cv->nchunk = (parts + cv->pc - 1) / cv->pc;
}

// Most tail chunks an object may use (the partial-sum ring is sized for it):
// chunks exist to give small channel counts enough workgroups, so beyond
// 4,096 chunk rows per object more of them only cost memory.
static int max_chunks_for(int C, int P) { return std::max(1, std::min(std::min(kMaxChunks, P - 2), 4096 / C)); }

// Shape of an object: uniform partitions of B over the first `head` taps
// (P = head / B), then (non-uniform objects) PB frame partitions of N = M * B.
struct ConvShape {
    int len = 0, B = 0, C = 0, M = 1, N = 0, P = 0, PB = 0, RB = 0, nchB = 0, pcB = 1, maxchunk = 1, shift = 0;
    int kp = 1;  // blocks per tail pass
    int nslots = 3;  // IR spectrum slots
    int kphase = 0;  // pass phase: passes in the launches of blocks b with (b + kphase) % kp == 0
    int irtype = IR_F32;
    bool td = false;  // time-domain object: IR slots and history in samples
};

static bool is_nu_method(uint32_t base)
{
    return base == TANX_METHOD_FFT_PARTITIONED_NONUNIFORM || base == TANX_METHOD_FHT_NONUNIFORM_PARTITIONED ||
           base == TANX_METHOD_FFT_NONUNIFORM_PARTITIONED;
}

// Frame chunks: enough workgroups at small channel counts, as for the tail.
static int frame_chunks(int C, int PB)
{
    if (PB <= 2) return 0;
    const int by_rows = (PB - 2 + 7) / 8;  // at least eight partitions per chunk
    return std::max(1, std::min(std::min(16, (1024 + C - 1) / C), by_rows));
}

// Modeled HBM bytes per block and channel of a frame multiple M (M = 1:
// uniform).  The reference picks its multiple from a CPU cost model of FFT
// and multiply-accumulate time (bestNUMultiple, ConvolutionImpl.cpp:201-246);
// on this device the multiply-accumulate streams HBM, so the model counts
// bytes: the first frame of taps as uniform partitions, the frame partitions
// 2..PB-1 once per frame (one slice per block), and the frame kernel's
// traffic spread over the frame's M blocks.
static double nu_model_bytes(int len, int B, int C, int M, int elem)
{
    const double row = 8.0 + elem;  // FDL + IR bytes per bin
    if (M == 1) return row * B * (len / B);
    const int N = M * B, PB = len / N - 1;
    const int nch = frame_chunks(C, PB);
    const double head = row * B * M;
    const double mac = PB > 2 ? row * B * (PB - 2) + 8.0 * B * nch : 0.0;
    const double frame = 2.0 * N * 4 + 2.0 * N * 8 + 2.0 * N * elem + (double)nch * N * 8 + 4.0 * N;
    return head + mac + frame / M;
}

static int choose_multiple(int len, int B, int C, int elem)
{
    int best = 1;
    double bytes = nu_model_bytes(len, B, C, 1, elem);
    for (int M = 2; M * B <= kMaxFrame && 2 * M * B <= len; M *= 2) {
        const double b = nu_model_bytes(len, B, C, M, elem);
        if (b < bytes) {
            bytes = b;
            best = M;
        }
    }
    return best;
}

static int32_t conv_shape(uint32_t method, uint32_t ir_len, uint32_t block, uint32_t channels,
                          const tanx_conv_options *opt, ConvShape &sh)
{
    const uint32_t prec = opt ? opt->ir_precision : TANX_IR_FP32;
    if (prec > TANX_IR_BF16) return TANX_INVALID_ARG;
    if (ir_len == 0 || block == 0 || channels == 0) return TANX_INVALID_ARG;
    if (!is_pow2(block) || block < 64 || block > 1024) return TANX_INVALID_ARG;
    uint32_t len = 1;
    while (len < ir_len) len <<= 1;
    if (len < block) len = block;
    if (len > (1u << 24)) return TANX_INVALID_ARG;
    sh.len = (int)len;
    sh.B = (int)block;
    sh.C = (int)channels;
    sh.irtype = prec == TANX_IR_FP16 ? IR_F16 : prec == TANX_IR_BF16 ? IR_BF16 : IR_F32;
    const uint32_t base = method & ~uint32_t(TANX_METHOD_USE_PROCESS_TAILTHREAD);
    if (base == TANX_METHOD_TIME_DOMAIN) {
        // direct convolution for short (HRTF-length) responses, FP32
        if (sh.len > kMaxTdLen || sh.irtype != IR_F32) return TANX_INVALID_ARG;
        if (opt && opt->ir_slots == 2) return TANX_NOT_SUPPORTED;
        sh.td = true;
        sh.P = 1;
        sh.maxchunk = 1;
        return TANX_OK;
    }
    int M = 1;
    if (is_nu_method(base)) {
        const uint32_t want = opt ? opt->nu_multiple : 0;
        if (want == 0) {
            M = choose_multiple(sh.len, sh.B, sh.C, ir_elem_bytes(sh.irtype));
        } else if (want > 1) {
            if (!is_pow2(want) || (uint64_t)want * block > (uint64_t)kMaxFrame || 2ull * want * block > len)
                return TANX_INVALID_ARG;
            M = (int)want;
        }
    }
    sh.M = M;
    if (M > 1) {
        // frames end where (block + 1 + shift) % M == 0: block M - 1 - shift
        // first (the frame clock runs `shift` blocks ahead of the block index)
        sh.shift = (int)((opt ? opt->nu_phase : 0) % (uint32_t)M);
        sh.N = M * sh.B;
        sh.P = M;
        sh.RB = sh.len / sh.N;
        sh.PB = sh.RB - 1;
        sh.nchB = frame_chunks(sh.C, sh.PB);
        sh.pcB = sh.nchB > 0 ? (sh.PB - 2 + sh.nchB - 1) / sh.nchB : 1;
        if (sh.nchB > 0) sh.nchB = (sh.PB - 2 + sh.pcB - 1) / sh.pcB;
    } else {
        sh.P = sh.len / sh.B;
    }
    sh.maxchunk = max_chunks_for(sh.C, sh.P);
    if (opt && opt->max_tail_chunks) sh.maxchunk = std::min(sh.maxchunk, (int)opt->max_tail_chunks);
    // multi-block tail passes: uniform objects outside finalize mode whose
    // IR has partitions enough to stream.  A pass reads every partition once
    // for kp blocks; the block that carries it streams the whole object (as
    // a tail every block would), the blocks between stream none, so the
    // worst block is no slower and the rows cost a quarter as much.  Objects
    // advanced together with pass phases 0..kp-1 take turns carrying their
    // pass, so no block streams every object's rows.
    const uint32_t want_kp = opt ? opt->pass_blocks : 0;
    if (want_kp > (uint32_t)kPassMax) return TANX_INVALID_ARG;
    const bool finalize = (method & TANX_METHOD_USE_PROCESS_TAILTHREAD) == TANX_METHOD_USE_PROCESS_FINALIZE;
    if (M == 1 && !finalize) {
        sh.kp = want_kp ? (int)want_kp : (sh.P >= 16 ? kPassDefault : 1);
        sh.kphase = (int)((opt ? opt->pass_phase : 0) % (uint32_t)sh.kp);
    }
    const uint32_t slots = opt ? opt->ir_slots : 0;
    if (slots != 0 && slots != 2 && slots != 3) return TANX_INVALID_ARG;
    if (slots == 2) {
        // non-uniform frames keep using the previous filter after a switch
        if (M != 1) return TANX_NOT_SUPPORTED;
        sh.nslots = 2;
    }
    return TANX_OK;
}

// Sizes of the pieces of an object's single device allocation, in order.
struct ConvLayout {
    size_t H, S, F, T, X, Ch, Run, HB, SB, FB, XR, AC, ZB;
    size_t total() const { return H + S + F + T + 3 * X + Ch + Run + HB + SB + FB + XR + AC + ZB; }
    size_t zeroed() const { return total() - Ch - Run; }
};

static ConvLayout conv_layout(const ConvShape &sh)
{
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t C = sh.C, B = sh.B, P = sh.P, N = sh.N, PB = sh.PB, elem = ir_elem_bytes(sh.irtype);
    ConvLayout l{};
    if (sh.td) {
        // [3][C][len] IR slots, [C][len] history (in F), I/O blocks
        l.H = al(3 * C * (size_t)sh.len * sizeof(float));
        l.F = al(C * (size_t)sh.len * sizeof(float));
        l.X = al(C * B * sizeof(float));
        l.Ch = al(C * sizeof(int));
        l.Run = al(C);
        return l;
    }
    l.H = al((size_t)sh.nslots * C * P * B * elem);
    l.S = al((size_t)sh.nslots * C * P * sizeof(float));
    l.F = al(C * P * B * sizeof(float2));
    l.T = al(4 * (size_t)sh.kp * C * (size_t)sh.maxchunk * B * sizeof(float2));
    l.X = al(C * B * sizeof(float));
    l.HB = al(3 * C * PB * N * elem);
    l.SB = al(3 * C * PB * sizeof(float));
    l.FB = al(C * (size_t)sh.RB * N * sizeof(float2));
    l.XR = al(C * 2 * N * sizeof(float));
    l.AC = al(4 * C * (size_t)sh.nchB * N * sizeof(float2));
    l.ZB = al(4 * C * N * sizeof(float));
    // channel list and run mask last: they are not cleared with the state
    l.Ch = al(C * sizeof(int));
    l.Run = al(C);
    return l;
}

extern "C" int32_t tanx_conv_estimate_bytes(uint32_t method, uint32_t ir_len, uint32_t block, uint32_t channels,
                                            const tanx_conv_options *opt, uint64_t *bytes)
{
    if (!bytes) return TANX_INVALID_POINTER;
    ConvShape sh;
    const int32_t rc = conv_shape(method, ir_len, block, channels, opt, sh);
    if (rc != TANX_OK) return rc;
    *bytes = conv_layout(sh).total() + sizeof(TickArgs) * kGraphBlocks;
    return TANX_OK;
}

extern "C" int32_t tanx_conv_create(tanx_ctx *ctx, uint32_t method, uint32_t ir_len,
                                    uint32_t block, uint32_t channels, tanx_conv **out)
{
    return tanx_conv_create_ex(ctx, method, ir_len, block, channels, nullptr, out);
}

extern "C" int32_t tanx_conv_create_ex(tanx_ctx *ctx, uint32_t method, uint32_t ir_len, uint32_t block,
                                       uint32_t channels, const tanx_conv_options *opt, tanx_conv **out)
{
    if (!ctx || !out) return TANX_INVALID_POINTER;
    *out = nullptr;
    const uint32_t base = method & ~uint32_t(TANX_METHOD_USE_PROCESS_TAILTHREAD);
    switch (base) {
    case TANX_METHOD_FFT_OVERLAP_ADD:
    case TANX_METHOD_FFT_PARTITIONED_UNIFORM:
    case TANX_METHOD_FFT_UNIFORM_PARTITIONED:
    case TANX_METHOD_FHT_UNIFORM_PARTITIONED:
    case TANX_METHOD_FFT_UNIFORM_HEAD_TAIL:
    case TANX_METHOD_FHT_UNIFORM_HEAD_TAIL:
    case TANX_METHOD_FFT_PARTITIONED_NONUNIFORM:
    case TANX_METHOD_FHT_NONUNIFORM_PARTITIONED:
    case TANX_METHOD_FFT_NONUNIFORM_PARTITIONED:
    case TANX_METHOD_TIME_DOMAIN:
        break;
    default:
        return TANX_NOT_SUPPORTED;
    }
    ConvShape sh;
    int32_t rc = conv_shape(method, ir_len, block, channels, opt, sh);
    if (rc != TANX_OK) return rc;

    DeviceGuard g(ctx->device);
    if (!g.ok) return TANX_NO_DEVICE;
    tanx_conv *cv = new (std::nothrow) tanx_conv();
    if (!cv) return TANX_OUT_OF_MEMORY;
    cv->ctx = ctx;
    ctx_retain(ctx);
    cv->method = method;
    cv->fmode = (method & TANX_METHOD_USE_PROCESS_TAILTHREAD) == TANX_METHOD_USE_PROCESS_FINALIZE;
    cv->irtype = sh.irtype;
    cv->L = (int)ir_len;
    cv->B = sh.B;
    cv->C = sh.C;
    cv->len = sh.len;
    cv->P = sh.P;
    cv->td = sh.td;
    cv->nu = sh.M > 1;
    cv->M = sh.M;
    cv->N = sh.N;
    cv->PB = sh.PB;
    cv->RB = sh.RB;
    cv->nchB = sh.nchB;
    cv->pcB = sh.pcB;
    cv->nu_shift = sh.shift;
    // partial-sum rows are only worth it at small channel counts
    cv->maxchunk = sh.maxchunk;
    set_chunks(cv, auto_chunks(cv->C, cv->P));
    cv->tail_nchunk = cv->nchunk;
    cv->kp = sh.kp;
    cv->kphase = sh.kphase;
    cv->nslots = sh.nslots;
    if (cv->nslots == 2) cv->upd = 0;  // slots 1 (current) and 0 (update)
    if (!cv->td) {
        cv->rows_bytes = (uint64_t)cv->C * cv->P * cv->B * sizeof(float2) + 2 * cv->slot_bytes() +
                         (cv->nu ? (uint64_t)cv->C * cv->RB * cv->N * sizeof(float2) + 2 * cv->bslot_bytes() : 0);
        ctx->pass_rows.fetch_add(cv->rows_bytes, std::memory_order_relaxed);
    }
    for (auto &r : cv->pchunk) r[0] = r[1] = cv->nchunk;
    // block 0: a silent history, its (zero) sums start at partition 2
    cv->covered = 0;
    cv->porg[0][0] = -1;

    const size_t C = cv->C, B = cv->B;
    const ConvLayout l = conv_layout(sh);
    const size_t bX = l.X;
    cv->dbytes = l.total();
    auto fail = [&](int32_t r) { conv_free(cv); return r; };
    hipError_t e = hipMalloc(&cv->dmem, cv->dbytes);
    if (e != hipSuccess) { cv->dmem = nullptr; return fail(map_hip(e)); }
    char *p = static_cast<char *>(cv->dmem);
    cv->H = p; p += l.H;
    cv->S = reinterpret_cast<float *>(p); p += l.S;
    cv->F = reinterpret_cast<float2 *>(p); p += l.F;
    cv->T = reinterpret_cast<float2 *>(p); p += l.T;
    cv->xprev = reinterpret_cast<float *>(p); p += bX;
    cv->din = reinterpret_cast<float *>(p); p += bX;
    cv->dout = reinterpret_cast<float *>(p); p += bX;
    cv->HBm = p; p += l.HB;
    cv->SBm = reinterpret_cast<float *>(p); p += l.SB;
    cv->FB = reinterpret_cast<float2 *>(p); p += l.FB;
    cv->XR = reinterpret_cast<float *>(p); p += l.XR;
    cv->AC = reinterpret_cast<float2 *>(p); p += l.AC;
    cv->ZB = reinterpret_cast<float *>(p); p += l.ZB;
    cv->dchan = reinterpret_cast<int *>(p); p += l.Ch;
    cv->drun = reinterpret_cast<uint8_t *>(p); p += l.Run;

    if (hipHostMalloc(&cv->hin, C * B * sizeof(float)) != hipSuccess ||
        hipHostMalloc(&cv->hout, C * B * sizeof(float)) != hipSuccess ||
        hipHostMalloc(&cv->hchan, C * sizeof(int)) != hipSuccess ||
        hipHostMalloc(&cv->hrun, kRunRing * C) != hipSuccess)
        return fail(TANX_OUT_OF_MEMORY);
    for (auto &ev : cv->run_ev)
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return fail(TANX_UNEXPECTED);
    if (hipMalloc(&cv->d_desc, sizeof(TickArgs) * kGraphBlocks) != hipSuccess ||
        hipHostMalloc(&cv->h_desc, sizeof(TickArgs) * kGraphBlocks * kDescRing) != hipSuccess)
        return fail(TANX_OUT_OF_MEMORY);
    for (auto &ev : cv->desc_ev)
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return fail(TANX_UNEXPECTED);
    for (auto &ev : cv->desc_ev)
        if (hipEventRecord(ev, ctx->proc) != hipSuccess) return fail(TANX_UNEXPECTED);
    if (hipEventCreateWithFlags(&cv->ev_upd, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&cv->ev_out, hipEventDisableTiming) != hipSuccess)
        return fail(TANX_UNEXPECTED);

    // all state starts at zero: the initial IR slot is silent
    // (ConvolutionImpl.cpp:115-117, 1640-1650)
    e = launch_zero(cv->dmem, l.zeroed(), ctx->proc);
    if (e == hipSuccess) {
        std::memset(cv->hrun, 1, C);
        e = hipMemcpyAsync(cv->drun, cv->hrun, C, hipMemcpyHostToDevice, ctx->proc);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->proc);
    if (e != hipSuccess) return fail(map_hip(e));

    cv->staged.assign(C, 0);
    cv->available.assign(C, 0);
    cv->flushed.assign(C, 0);
    cv->run_now.assign(C, 1);
    cv->run_dev.assign(C, 1);
    cv->tail_left.assign(C, 0);
    *out = cv;
    return TANX_OK;
}

extern "C" int32_t tanx_conv_destroy(tanx_conv *cv)
{
    if (!cv) return TANX_INVALID_POINTER;
    conv_free(cv);
    return TANX_OK;
}

// zero one channel's history: FDL ring, previous block, tail partial sums
static int32_t flush_channel(tanx_conv *cv, int ch)
{
    hipStream_t s = cv->ctx->proc;
    const size_t B = cv->B;
    if (cv->td) return map_hip(launch_zero(cv->td_hist() + (size_t)ch * cv->len, (size_t)cv->len * sizeof(float), s));
    hipError_t e = launch_zero(cv->F + (size_t)ch * cv->P * B, (size_t)cv->P * B * sizeof(float2), s);
    if (e == hipSuccess) e = launch_zero(cv->xprev + (size_t)ch * B, B * sizeof(float), s);
    // partial sums are laid out [C][tail_nchunk][B]: the next head reads them
    if (cv->kp > 1) {
        // every ring slot: sums computed ahead hold the history being dropped
        for (int r = 0; r < 2 * cv->kp; r++)
            for (int f = 0; f < 2 && e == hipSuccess; f++)
                e = launch_zero(cv->tail_slot(r, f) + (size_t)ch * cv->pchunk[r][f] * B,
                                (size_t)cv->pchunk[r][f] * B * sizeof(float2), s);
    } else {
        for (int f = 0; f < 2 && e == hipSuccess; f++)
            e = launch_zero(cv->tail_ptr(cv->tread, f) + (size_t)ch * cv->tail_nchunk * B,
                            (size_t)cv->tail_nchunk * B * sizeof(float2), s);
    }
    if (cv->nu) {
        // frame ring, input ring, frame partials and outputs of the channel
        const size_t N = cv->N;
        if (e == hipSuccess) e = launch_zero(cv->FB + (size_t)ch * cv->RB * N, (size_t)cv->RB * N * sizeof(float2), s);
        if (e == hipSuccess) e = launch_zero(cv->XR + (size_t)ch * 2 * N, 2 * N * sizeof(float), s);
        for (int par = 0; par < 2; par++)
            for (int f = 0; f < 2; f++) {
                if (e == hipSuccess && cv->nchB > 0)
                    e = launch_zero(cv->acc_ptr(par, f) + (size_t)ch * cv->nchB * N,
                                    (size_t)cv->nchB * N * sizeof(float2), s);
                if (e == hipSuccess) e = launch_zero(cv->zb_ptr(par, f) + (size_t)ch * N, N * sizeof(float), s);
            }
    }
    return map_hip(e);
}

// mirrors the channel flag bookkeeping of ProcessInternal
// (ConvolutionImpl.cpp:2868-2907) and returns the running count.  A channel
// that returns from the stopped (available) state starts from silence: the
// reference flushes a draining channel before its last drain block and would
// otherwise resume with that block's state.
static int32_t channel_flags(tanx_conv *cv, const uint32_t *flags, size_t n, int *nrun)
{
    int running = 0;
    for (int ch = 0; ch < cv->C; ch++) {
        bool do_flush = false;
        const bool was_available = cv->available[ch] != 0;
        if (!flags) {
            cv->available[ch] = 0;
            cv->tail_left[ch] = cv->len;
        } else if (flags[ch] & TANX_CHANNEL_FLAG_STOP_INPUT) {
            if (flags[ch] & TANX_CHANNEL_FLAG_FLUSH_STREAM) {
                cv->available[ch] = 1;
                cv->tail_left[ch] = 0;
            } else if (cv->tail_left[ch] > 0) {
                cv->available[ch] = 0;
                cv->tail_left[ch] -= (int)n;
                if (cv->tail_left[ch] <= 0) do_flush = true;
            } else {
                cv->available[ch] = 1;
                cv->tail_left[ch] = 0;
            }
        } else {
            cv->available[ch] = 0;
            cv->tail_left[ch] = cv->len;
        }
        if (flags && (flags[ch] & TANX_CHANNEL_FLAG_FLUSH_STREAM)) {
            if (!cv->flushed[ch]) do_flush = true;
            cv->flushed[ch] = 1;
        } else {
            cv->flushed[ch] = 0;
        }
        if (was_available && !cv->available[ch]) do_flush = true;
        if (do_flush) {
            int32_t rc = flush_channel(cv, ch);
            if (rc != TANX_OK) return rc;
        }
        cv->run_now[ch] = cv->available[ch] ? 0 : 1;
        running += cv->run_now[ch];
    }
    *nrun = running;
    return TANX_OK;
}

static int32_t sync_run_mask(tanx_conv *cv)
{
    if (cv->run_now == cv->run_dev) return TANX_OK;
    const int slot = cv->run_ring;
    cv->run_ring = (cv->run_ring + 1) % kRunRing;
    TANX_CHECK(hipEventSynchronize(cv->run_ev[slot]));  // pinned entry free again
    uint8_t *h = cv->hrun + (size_t)slot * cv->C;
    std::memcpy(h, cv->run_now.data(), cv->C);
    TANX_CHECK(hipMemcpyAsync(cv->drun, h, cv->C, hipMemcpyHostToDevice, cv->ctx->proc));
    TANX_CHECK(hipEventRecord(cv->run_ev[slot], cv->ctx->proc));
    cv->run_dev = cv->run_now;
    return TANX_OK;
}

// ordinary row loads for this object's passes: the context's live pass rows fit the cached bound
static bool cached_rows(const tanx_conv *cv)
{
    return !cv->td && cv->ctx->pass_rows.load(std::memory_order_relaxed) <= kCachedRowsMax;
}

static TickArgs tick_common(tanx_conv *cv)
{
    TickArgs a{};
    a.B = cv->B;
    a.C = cv->C;
    a.P = cv->P;
    a.irtype = cv->irtype;
    a.run = cv->drun;
    const int lB = __builtin_ctz((unsigned)cv->B);
    a.twM = tw_table(cv->ctx->tw, lB);
    a.twN = tw_table(cv->ctx->tw, lB + 1);
    a.fdl = cv->F;
    a.cached = cached_rows(cv) ? 1 : 0;
    a.kp = cv->kp;
    a.kn = 1;
    a.kgap = 1;
    a.dA = a.dB = 1;
    return a;
}

// Tail role arguments for the block after the current one: partitions
// 2..P-1, with both filters when that block is the crossfade block.  With
// multi-block passes, the pass (if any) of a launch whose heads are the
// current block alone.
static int fill_pass(tanx_conv *cv, long long w, long long e, TickArgs &a);

static int fill_tail(tanx_conv *cv, TickArgs &a)
{
    if (cv->kp > 1) {
        const long long e = (long long)cv->ticks;
        return std::max(1, fill_pass(cv, e - 1, e, a));
    }
    // cv->ticks is the index of the block this tail is for
    const int nf = (cv->delayed && cv->ticks == cv->switch_at) ? 2 : 1;
    a.nchunk = cv->nchunk;
    a.pc = cv->pc;
    a.tbase = cv->pos;  // pos already advanced past the current block
    a.tHa = cv->slot_ptr(cv->cur);
    a.tHb = nf == 2 ? cv->slot_ptr(cv->upd) : nullptr;
    a.tSa = cv->scale_ptr(cv->cur);
    a.tSb = nf == 2 ? cv->scale_ptr(cv->upd) : nullptr;
    a.tTa = cv->tail_ptr(cv->tread ^ 1, 0);
    a.tTb = nf == 2 ? cv->tail_ptr(cv->tread ^ 1, 1) : nullptr;
    return nf;
}

// Pass of a launch running the heads of blocks [w, e) (cv->ticks == e, the
// switch state that after block e-1): when block e has no partial sums yet,
// the sums of blocks e .. c (c the next block on the pass grid) from FDL
// rows before w (output k: partitions >= e - w + k + 1), with the filter
// each block will use: the old one before an armed switch, both at its
// crossfade block, the new one after it.  When a switch armed at block e-1
// falls in blocks an earlier pass covered, the old filter's sums of those
// blocks stay valid and a catch-up pass computes the new filter's alone:
// the crossfade block's second sums and the sums of the blocks after it.
// Returns the number of filters the pass reads (0: no pass).
static int fill_pass(tanx_conv *cv, long long w, long long e, TickArgs &a)
{
    const long long sw = cv->delayed ? (long long)cv->switch_at : -1;
    const bool stale = sw >= e && sw <= cv->covered;
    a.tHb = nullptr;
    a.tSb = nullptr;
    for (int k = 0; k < kPassMax; k++) a.kTA[k] = a.kTB[k] = nullptr;
    a.tTa = a.tTb = nullptr;
    if (cv->covered >= e && !stale) {
        a.nchunk = 0;
        a.kn = 1;
        return 0;
    }
    // the pass ends on the phase grid: its last block c satisfies
    // (c + kphase) % kp == 0, so the next pass runs in that block's launch
    const int kn = (int)(((-(long long)cv->kphase - e) % cv->kp + cv->kp) % cv->kp) + 1;
    const long long s_rel = (sw >= e && sw < e + kn) ? sw - e : -1;
    a.nchunk = cv->nchunk;
    a.pc = cv->pc;
    a.kn = kn;
    a.kgap = (int)(e - w);
    a.tbase = cv->pos;  // ring slot of block e
    // an earlier pass's grid ends at covered, so this pass ends there too;
    // its old-filter sums of the crossfade block are kept when they were
    // computed with the current chunking (the head reads both with one)
    if (stale && e + kn - 1 == cv->covered && cv->pchunk[cv->pring(sw)][0] == cv->nchunk) {
        a.tHa = cv->slot_ptr(cv->upd);
        a.tSa = cv->scale_ptr(cv->upd);
        for (int k = (int)s_rel; k < kn; k++) {
            const long long blk = e + k;
            const int r = cv->pring(blk);
            const int f = k == s_rel ? 1 : 0;  // after the switch the new filter is the block's own
            a.kTA[k] = cv->tail_slot(blk, f);
            cv->porg[r][f] = w;
            cv->pchunk[r][f] = cv->nchunk;
        }
        return 1;
    }
    a.tHa = cv->slot_ptr(cv->cur);
    a.tSa = cv->scale_ptr(cv->cur);
    bool second = false;
    for (int k = 0; k < kn; k++) {
        const long long blk = e + k;
        const int r = cv->pring(blk);
        float2 *own = cv->tail_slot(blk, 0);
        if (s_rel < 0 || k < s_rel) {
            a.kTA[k] = own;
        } else if (k == s_rel) {
            a.kTA[k] = own;
            a.kTB[k] = cv->tail_slot(blk, 1);
            cv->porg[r][1] = w;
            cv->pchunk[r][1] = cv->nchunk;
            second = true;
        } else {
            a.kTB[k] = own;  // after the switch the new filter is the block's own
            second = true;
        }
        cv->porg[r][0] = w;
        cv->pchunk[r][0] = cv->nchunk;
    }
    if (second) {
        a.tHb = cv->slot_ptr(cv->upd);
        a.tSb = cv->scale_ptr(cv->upd);
    }
    cv->covered = e + kn - 1;
    return second ? 2 : 1;
}

// After a tail has been enqueued: the next head reads its buffer, and the
// update hand-off the reference's update thread performs
// (ConvolutionImpl.cpp:1046-1079, 2004-2199).
static void tail_done(tanx_conv *cv)
{
    cv->tread ^= 1;
    cv->tail_nchunk = cv->nchunk;
    cv->tail_due = false;
    if (!cv->crossfading && cv->pending) {
        cv->pending = false;
        cv->ready = true;
        std::fill(cv->staged.begin(), cv->staged.end(), 0);
    }
}

static int32_t launch_timed(tanx_conv *cv, KernelTimer &kt, int tail_nf, const TickArgs &a,
                            uint64_t alg_bytes, uint64_t &alg_sum, const FrameArgs *fa = nullptr)
{
    hipStream_t s = cv->ctx->proc;
    if (cv->timing) {
        // the launches stamp the events themselves: the time of the kernels alone
        hipEvent_t e0 = nullptr, e1 = nullptr;
        TANX_CHECK(kt.acquire(&e0, &e1));
        TANX_CHECK(launch_tick(cv->B, tail_nf, a, s, e0, fa ? nullptr : e1));
        if (fa) TANX_CHECK(launch_frame(cv->irtype, *fa, s, e1));
        kt.commit();
        alg_sum += alg_bytes;
        if (a.nhead > 0) cv->timed_blocks++;
        return TANX_OK;
    }
    TANX_CHECK(launch_tick(cv->B, tail_nf, a, s));
    if (fa) TANX_CHECK(launch_frame(cv->irtype, *fa, s));
    return TANX_OK;
}

// algorithmic bytes of the roles (regular blocks count one filter)
static uint64_t head_bytes(const tanx_conv *cv, int nf)
{
    // x_{t-1}, x_t, y, the new FDL row, the previous FDL row, and per filter
    // partitions 0 and 1
    return (uint64_t)cv->C * cv->B * (3 * sizeof(float) + 2 * sizeof(float2) + 2 * ir_elem_bytes(cv->irtype) * nf);
}
static uint64_t tail_bytes(const tanx_conv *cv, int nf)
{
    // FDL rows and nf IR rows (plus their scales) of partitions 2..P-1
    const uint64_t parts = cv->P > 2 ? (uint64_t)(cv->P - 2) : 0;
    const uint64_t scales = cv->irtype == IR_F32 ? 0 : sizeof(float) * nf;
    return (uint64_t)cv->C * parts * (cv->B * (sizeof(float2) + (uint64_t)ir_elem_bytes(cv->irtype) * nf) + scales);
}

// partitions 2..d the head adds beyond partitions 0 and 1 (multi-block
// passes): an IR row and an FDL row each
static uint64_t near_bytes(const tanx_conv *cv, const TickArgs &a)
{
    const int extra = std::max(0, a.dA - 1) + (a.xfade ? std::max(0, a.dB - 1) : 0);
    return (uint64_t)extra * cv->C * cv->B * (sizeof(float2) + (uint64_t)ir_elem_bytes(cv->irtype));
}

// frame roles of a non-uniform block: the multiply-accumulate slice (nf
// filters) and, at the end of a frame, the frame kernel
static uint64_t frame_mac_bytes(const tanx_conv *cv, int nf)
{
    if (!cv->nu || cv->nchB == 0) return 0;
    const uint64_t scales = cv->irtype == IR_F32 ? 0 : sizeof(float) * nf;
    return (uint64_t)cv->C * (cv->PB - 2) * (cv->B * (sizeof(float2) + (uint64_t)ir_elem_bytes(cv->irtype) * nf) + scales) +
           (uint64_t)cv->C * cv->nchB * cv->B * sizeof(float2) * nf;
}
static uint64_t frame_kernel_bytes(const tanx_conv *cv, int nf)
{
    // 2N input samples, X_g written, X_{g-1} read, per filter two spectrum
    // rows, the partials and the N output samples
    const uint64_t N = cv->N;
    return (uint64_t)cv->C * (2 * N * sizeof(float) + 2 * N * sizeof(float2) +
                              nf * (2 * N * (uint64_t)ir_elem_bytes(cv->irtype) + cv->nchB * N * sizeof(float2) +
                                    N * sizeof(float)));
}

// ProcessFinalize: the tail for the next block as its own launch.
static int32_t enqueue_tail(tanx_conv *cv)
{
    if (!cv->tail_due) return TANX_OK;
    TickArgs a = tick_common(cv);
    a.nhead = 0;
    const int nf = fill_tail(cv, a);
    if (a.nchunk > 0) {
        int32_t rc = launch_timed(cv, cv->t_tail, nf, a, tail_bytes(cv, nf), cv->tail_alg_bytes);
        if (rc != TANX_OK) return rc;
    }
    tail_done(cv);
    return TANX_OK;
}

// Process state machine (ConvolutionImpl.cpp:848-1043) for one block: flag
// bookkeeping, switch schedule, slot rotation, and the launch arguments of
// the block's head (and, with_tail, the next block's tail).  Sets *tail_nf.
// Switch schedule of block b (ConvolutionImpl.cpp:869-927): arms a handed-off
// update, and at the armed block crossfades and rotates the slots (current ->
// previous, update -> current).  Returns whether block b crossfades.
static bool advance_switch(tanx_conv *cv, uint64_t b)
{
    bool xfade = false;
    if (!cv->pending && cv->ready && !cv->delayed) {
        cv->ready = false;
        cv->delayed = true;
        // uniform: the next block; non-uniform: a frame boundary late enough
        // for the armed filter's frame partials (M + 1 blocks ahead)
        const uint64_t M = (uint64_t)cv->M, d = (uint64_t)cv->nu_shift;
        cv->switch_at = cv->nu ? ((b + d + M + 1 + M - 1) / M) * M - d : b + 1;
    } else if (cv->delayed && b == cv->switch_at) {
        xfade = true;
    }
    if (xfade) {
        // the armed filter becomes current; the update slot is the one no
        // filter uses after this block (with 2 slots: the previous filter's,
        // which the crossfade block itself still reads, so an update is
        // written only after this block in stream order)
        const int old = cv->cur;
        cv->prev = old;
        cv->cur = cv->upd;
        cv->upd = cv->nslots == 3 ? (cv->upd + 1) % 3 : old;
    }
    return xfade;
}

// Frame roles of block b of a non-uniform object (after this block's slot
// rotation).  The slice for frame gm = (b+1)/M feeds blocks (gm+1)M.. and
// uses the filters in effect there: both while a switch at (gm+1)M is
// armed, the armed one alone at the block before an armed switch (gm's
// blocks all follow it), else the current one.  The frame kernel at the last
// block of frame g feeds block b+1.
static void nu_block_args(tanx_conv *cv, uint64_t b_obj, bool xfade, TickArgs &a, FrameArgs &fa, bool *frame_due)
{
    const uint64_t M = (uint64_t)cv->M;
    // the object's frame clock (a frame phase shifts it); b and the switch
    // block are read in it below, ring slots and parities follow it
    const uint64_t b = b_obj + (uint64_t)cv->nu_shift;
    const uint64_t switch_at = cv->switch_at + (uint64_t)cv->nu_shift;
    a.nu = 1;
    a.M = cv->M;
    a.N = cv->N;
    a.xring = cv->XR;
    a.xslot = (int)(b % (2 * M));
    a.zoff = (int)(b % M) * cv->B;
    const long long fo = (long long)(b / M) - 1;  // frame whose output covers block b
    a.zA = nullptr;
    a.zB = nullptr;
    if (fo >= 0) {
        const int par = (int)(fo & 1);
        a.zA = cv->zb_ptr(par, (!xfade && fo == cv->zsel_frame) ? 1 : 0);
        a.zB = xfade ? cv->zb_ptr(par, 1) : nullptr;
    }
    a.PB = cv->PB;
    a.RB = cv->RB;
    a.fb = cv->FB;
    if (cv->nchB > 0) {
        const uint64_t gm = (b + 1) / M;
        a.nchB = cv->nchB;
        a.pcB = cv->pcB;
        a.fbase = (int)(gm & (uint64_t)(cv->RB - 1));
        a.bslice = (int)((b + 1) % M);
        const bool both = cv->delayed && switch_at == (gm + 1) * M;
        const bool next_only = !both && cv->delayed && switch_at == b + 1;
        const int fa0 = next_only ? cv->upd : cv->cur;
        a.bHa = cv->bslot_ptr(fa0);
        a.bSa = cv->bscale_ptr(fa0);
        a.bHb = both ? cv->bslot_ptr(cv->upd) : nullptr;
        a.bSb = both ? cv->bscale_ptr(cv->upd) : nullptr;
        a.bTa = cv->acc_ptr((int)(gm & 1), 0);
        a.bTb = both ? cv->acc_ptr((int)(gm & 1), 1) : nullptr;
    }
    *frame_due = ((b + 1) % M) == 0;
    if (*frame_due) {
        const uint64_t g = (b + 1) / M - 1;
        const bool two = cv->delayed && switch_at == b + 1;
        fa = FrameArgs{};
        fa.B = cv->B;
        fa.M = cv->M;
        fa.N = cv->N;
        fa.C = cv->C;
        fa.PB = cv->PB;
        fa.RB = cv->RB;
        fa.run = cv->drun;
        fa.xring = cv->XR;
        fa.xlast = (int)(b % (2 * M));
        fa.fb = cv->FB;
        fa.frow = (int)(g & (uint64_t)(cv->RB - 1));
        fa.Ha = cv->bslot_ptr(cv->cur);
        fa.Sa = cv->bscale_ptr(cv->cur);
        fa.Hb = two ? cv->bslot_ptr(cv->upd) : nullptr;
        fa.Sb = two ? cv->bscale_ptr(cv->upd) : nullptr;
        fa.nch = cv->nchB;
        fa.Ta = cv->nchB ? cv->acc_ptr((int)(g & 1), 0) : nullptr;
        fa.Tb = (two && cv->nchB) ? cv->acc_ptr((int)(g & 1), 1) : nullptr;
        fa.za = cv->zb_ptr((int)(g & 1), 0);
        fa.zb = two ? cv->zb_ptr((int)(g & 1), 1) : nullptr;
        const int lN = __builtin_ctz((unsigned)cv->N);
        fa.twN = tw_table(cv->ctx->tw, lN);
        fa.tw2N = tw_table(cv->ctx->tw, lN + 1);
        if (two) cv->zsel_frame = (long long)g;
    }
}

static int32_t prepare_block(tanx_conv *cv, const float *din, long long in_stride, float *dout,
                             long long out_stride, size_t n, const uint32_t *flags, bool with_tail,
                             TickArgs &a, int *tail_nf, bool *xfade_out,
                             FrameArgs *fa = nullptr, bool *frame_due = nullptr)
{
    int nrun = 0;
    int32_t rc = channel_flags(cv, flags, n, &nrun);
    if (rc != TANX_OK) return rc;
    if (nrun == 0) return TANX_WRONG_STATE;
    rc = sync_run_mask(cv);
    if (rc != TANX_OK) return rc;

    const uint64_t b = cv->ticks;
    const bool xfade = advance_switch(cv, b);
    a = tick_common(cv);
    a.nhead = cv->C;
    a.xin = din;
    a.in_stride = in_stride;
    a.yout = dout;
    a.out_stride = out_stride;
    a.xprev = cv->xprev;
    a.slot = cv->pos;
    a.Ha = cv->slot_ptr(xfade ? cv->prev : cv->cur);
    a.Hb = xfade ? cv->slot_ptr(cv->cur) : nullptr;
    a.Sa = cv->scale_ptr(xfade ? cv->prev : cv->cur);
    a.Sb = xfade ? cv->scale_ptr(cv->cur) : nullptr;
    a.Ta = cv->tail_ptr(cv->tread, 0);
    a.Tb = xfade ? cv->tail_ptr(cv->tread, 1) : cv->tail_ptr(cv->tread, 0);
    a.tchunk = cv->tail_nchunk;
    a.xfade = xfade ? 1 : 0;
    if (cv->kp > 1) {
        // this block's partial sums, and the partitions its head adds: those
        // after the launch whose pass computed the sums
        const long long bb = (long long)b;
        const int r = cv->pring(bb);
        if (bb > cv->covered) return TANX_UNEXPECTED;  // a launch before it computes its sums
        a.Ta = cv->tail_slot(bb, 0);
        a.Tb = xfade ? cv->tail_slot(bb, 1) : a.Ta;
        a.dA = (int)(bb - cv->porg[r][0]);
        a.dB = xfade ? (int)(bb - cv->porg[r][1]) : 1;
        a.tchunk = cv->pchunk[r][0];
        const int near = kHeadBatch + pass_km(cv->kp) - 1;  // what a batched head adds, at most
        if (a.dA < 1 || a.dA > near || a.dB < 1 || a.dB > near ||
            (xfade && cv->pchunk[r][1] != a.tchunk))
            return TANX_UNEXPECTED;
    }

    cv->pos = (cv->pos + 1) % cv->P;
    cv->crossfading = xfade;
    cv->last_xfade = xfade;
    if (xfade) {
        cv->delayed = false;
        cv->xfades++;
    }
    bool fdue = false;
    if (cv->nu) {
        FrameArgs tmp;
        nu_block_args(cv, b, xfade, a, fa ? *fa : tmp, &fdue);
    }
    if (frame_due) *frame_due = fdue;
    cv->ticks++;
    cv->tail_due = true;

    int nf = 1;
    if (with_tail) nf = fill_tail(cv, a);
    *tail_nf = nf;
    *xfade_out = xfade;
    return TANX_OK;
}

static void td_common(const tanx_conv *cv, TdArgs &t)
{
    t = TdArgs{};
    t.B = cv->B;
    t.C = cv->C;
    t.len = cv->len;
    t.run = cv->drun;
    t.hist = cv->td_hist();
    t.h_stride = cv->len;
}

// One block of a time-domain object: the same switch schedule, one launch.
static int32_t process_td_block(tanx_conv *cv, const float *din, long long in_stride, float *dout,
                                long long out_stride, size_t n, const uint32_t *flags)
{
    int nrun = 0;
    int32_t rc = channel_flags(cv, flags, n, &nrun);
    if (rc != TANX_OK) return rc;
    if (nrun == 0) return TANX_WRONG_STATE;
    rc = sync_run_mask(cv);
    if (rc != TANX_OK) return rc;
    const bool xfade = advance_switch(cv, cv->ticks);
    TdArgs t;
    td_common(cv, t);
    t.n = cv->B;
    t.xin = din;
    t.in_stride = in_stride;
    t.yout = dout;
    t.out_stride = out_stride;
    t.pos = cv->tdpos;
    t.Ha = cv->td_slot(xfade ? cv->prev : cv->cur);
    t.Hb = xfade ? cv->td_slot(cv->cur) : nullptr;
    t.xfade = xfade ? 1 : 0;
    hipStream_t s = cv->ctx->proc;
    KernelTimer &kt = cv->t_head;
    if (cv->timing) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        TANX_CHECK(kt.acquire(&e0, &e1));
        TANX_CHECK(launch_td(t, s, e0, e1));
        kt.commit();
        // history window, filters, input and output
        cv->head_alg_bytes += (uint64_t)cv->C * (cv->len * sizeof(float) * (xfade ? 3 : 2) + 3 * cv->B * sizeof(float));
        cv->timed_blocks++;
    } else {
        TANX_CHECK(launch_td(t, s));
    }
    cv->tdpos = (cv->tdpos + cv->B) & (cv->len - 1);
    cv->crossfading = xfade;
    cv->last_xfade = xfade;
    if (xfade) {
        cv->delayed = false;
        cv->xfades++;
    }
    cv->ticks++;
    cv->tail_due = true;
    tail_done(cv);  // no tail: hand off a staged update now
    return TANX_OK;
}

// Blocks of an object with multi-block passes prepared but not launched
// yet: consecutive heads of one launch (flush_heads), within one Process
// call.
struct HeadQueue {
    TickArgs a[kHeadBatch];
    int n = 0;
};

// Launches the queued heads [w, e) (e = cv->ticks) with the pass fill_pass
// chooses: one head with k_tick (head role, pass role), several with
// k_window (batched heads, pass role) or k_heads (no pass).
static int32_t flush_heads(tanx_conv *cv, HeadQueue &q)
{
    const int n = q.n;
    if (n == 0) return TANX_OK;
    q.n = 0;
    const long long e = (long long)cv->ticks, w = e - n;
    TickArgs p = q.a[0];
    const int pnf = fill_pass(cv, w, e, p);
    uint64_t hb = 0;
    for (int j = 0; j < n; j++) hb += head_bytes(cv, q.a[j].xfade ? 2 : 1) + near_bytes(cv, q.a[j]);
    const uint64_t tb = pnf > 0 ? tail_bytes(cv, pnf) : 0;
    KernelTimer &kt = pnf > 0 ? cv->t_tail : cv->t_head;
    uint64_t &alg = pnf > 0 ? cv->tail_alg_bytes : cv->head_alg_bytes;
    if (n == 1) return launch_timed(cv, kt, std::max(1, pnf), p, hb + tb, alg);
    const TickArgs &a0 = q.a[0];
    HeadBatchArgs h{};
    h.C = a0.C;
    h.P = a0.P;
    h.run = a0.run;
    h.twM = a0.twM;
    h.twN = a0.twN;
    h.fdl = a0.fdl;
    h.xprev = a0.xprev;
    h.in_stride = a0.in_stride;
    h.out_stride = a0.out_stride;
    h.n = n;
    h.kp = cv->kp;
    // entries past the batch repeat its last block (the kernel loads them unused)
    for (int j = 0; j < kHeadBatch; j++) {
        const TickArgs &a = q.a[std::min(j, n - 1)];
        // one filter for the whole batch (a crossfade block runs alone)
        if (a.xfade || a.dA > kHeadBatch + pass_km(cv->kp) - 1 || a.Ha != a0.Ha || a.Sa != a0.Sa)
            return TANX_UNEXPECTED;
        h.xin[j] = a.xin;
        h.yout[j] = a.yout;
        h.H[j] = a.Ha;
        h.S[j] = a.Sa;
        h.T[j] = a.Ta;
        h.slot[j] = a.slot;
        h.tchunk[j] = a.tchunk;
        h.d[j] = a.dA;
    }
    hipStream_t s = cv->ctx->proc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (cv->timing) {
        TANX_CHECK(kt.acquire(&e0, &e1));
    }
    if (pnf > 0) TANX_CHECK(launch_window(cv->B, pnf, h, p, s, e0, e1));
    else TANX_CHECK(launch_heads(cv->B, cv->irtype, h, s, e0, e1));
    if (cv->timing) {
        kt.commit();
        alg += hb + tb;
        cv->timed_blocks += (uint64_t)n;
    }
    return TANX_OK;
}

// One block of the fused device path of an object with multi-block passes:
// joins the queued heads unless it must start a launch (no partial sums
// yet, a crossfade block, channel flags, a full queue), and ends the
// launch when it crossfades or arms a switch (its pass then computes the
// new filter's sums) or when it must run alone.
static int32_t process_block_kp(tanx_conv *cv, const float *din, long long in_stride, float *dout,
                                long long out_stride, size_t n, const uint32_t *flags, HeadQueue &q)
{
    const long long b = (long long)cv->ticks;
    const bool xf_next = cv->delayed && (long long)cv->switch_at == b;
    if (q.n > 0 && (b > cv->covered || xf_next || flags || q.n == kHeadBatch)) {
        const int32_t rc = flush_heads(cv, q);
        if (rc != TANX_OK) return rc;
    }
    TickArgs a;
    int nf = 1;
    bool xf = false;
    const int32_t rc = prepare_block(cv, din, in_stride, dout, out_stride, n, flags, false, a, &nf, &xf);
    if (rc != TANX_OK) {
        // the blocks already advanced are launched before the error returns
        const int32_t frc = flush_heads(cv, q);
        return frc != TANX_OK ? frc : rc;
    }
    q.a[q.n++] = a;
    tail_done(cv);
    const bool armed = cv->delayed && (long long)cv->switch_at == b + 1;
    if (xf || armed || flags || q.n == kHeadBatch || cv->B > kHeadBatchMaxB) return flush_heads(cv, q);
    return TANX_OK;
}

// One block launched directly; with_tail also runs the next block's tail in
// the same launch.
static int32_t process_block(tanx_conv *cv, const float *din, long long in_stride, float *dout,
                             long long out_stride, size_t n, const uint32_t *flags, bool with_tail)
{
    if (cv->td) return process_td_block(cv, din, in_stride, dout, out_stride, n, flags);
    TickArgs a;
    FrameArgs fa;
    int tail_nf = 1;
    bool xfade = false, frame_due = false;
    int32_t rc = prepare_block(cv, din, in_stride, dout, out_stride, n, flags, with_tail, a, &tail_nf,
                               &xfade, &fa, &frame_due);
    if (rc != TANX_OK) return rc;
    uint64_t hb = head_bytes(cv, xfade ? 2 : 1) + frame_mac_bytes(cv, a.bHb ? 2 : 1) + near_bytes(cv, a);
    if (frame_due) hb += frame_kernel_bytes(cv, fa.Hb ? 2 : 1);
    const FrameArgs *fp = frame_due ? &fa : nullptr;
    if (with_tail) {
        const uint64_t tb = a.nchunk > 0 ? tail_bytes(cv, tail_nf) : 0;
        // with multi-block passes, the launches that run a pass are timed as
        // tail launches and the head-only ones as head launches
        const bool pass = cv->kp > 1 && a.nchunk > 0;
        rc = launch_timed(cv, pass ? cv->t_tail : cv->t_head, tail_nf, a, hb + tb,
                          pass ? cv->tail_alg_bytes : cv->head_alg_bytes, fp);
        if (rc != TANX_OK) return rc;
        tail_done(cv);
    } else {
        rc = launch_timed(cv, cv->t_head, 1, a, hb, cv->head_alg_bytes, fp);
        if (rc != TANX_OK) return rc;
    }
    return TANX_OK;
}

// Instantiated graph of k descriptor-driven tick launches (one per block).
static int32_t graph_for(tanx_conv *cv, int k, hipGraphExec_t *out)
{
    const unsigned blocks = (unsigned)(cv->C + cv->C * cv->nchunk);
    for (auto &g : cv->graphs)
        if (g.k == k && g.blocks == blocks) {
            *out = g.exec;
            return TANX_OK;
        }
    hipStream_t s = cv->ctx->proc;
    hipGraph_t graph = nullptr;
    TANX_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int j = 0; j < k && e == hipSuccess; j++) e = launch_tick_desc(cv->B, cv->irtype, cv->d_desc, j, blocks, s);
    hipError_t e2 = hipStreamEndCapture(s, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return map_hip(e != hipSuccess ? e : e2);
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return map_hip(e);
    cv->graphs.push_back({k, blocks, exec});
    *out = exec;
    return TANX_OK;
}

// Device-memory blocks [0, nb) replayed through graphs of up to
// kGraphBlocks launches: the host runs the state machine for the batch,
// copies the batch's launch descriptors to the device, and launches the graph.
static int32_t process_device_graph(tanx_conv *cv, const float *d_in, size_t in_stride, float *d_out,
                                    size_t out_stride, size_t nb, size_t *processed)
{
    hipStream_t s = cv->ctx->proc;
    const size_t B = cv->B;
    size_t b = 0;
    while (b < nb) {
        const int k = (int)std::min<size_t>(kGraphBlocks, nb - b);
        const int slot = cv->desc_ring;
        cv->desc_ring = (cv->desc_ring + 1) % kDescRing;
        TANX_CHECK(hipEventSynchronize(cv->desc_ev[slot]));  // pinned entry free
        TickArgs *h = cv->h_desc + (size_t)slot * kGraphBlocks;
        for (int j = 0; j < k; j++) {
            int nf = 1;
            bool xf = false;
            int32_t rc = prepare_block(cv, d_in + (b + j) * B, (long long)in_stride, d_out + (b + j) * B,
                                       (long long)out_stride, B, nullptr, true, h[j], &nf, &xf);
            if (rc != TANX_OK) return rc;
            tail_done(cv);
        }
        hipGraphExec_t exec = nullptr;
        int32_t rc = graph_for(cv, k, &exec);
        if (rc != TANX_OK) return rc;
        TANX_CHECK(hipMemcpyAsync(cv->d_desc, h, sizeof(TickArgs) * k, hipMemcpyHostToDevice, s));
        TANX_CHECK(hipEventRecord(cv->desc_ev[slot], s));
        TANX_CHECK(hipGraphLaunch(exec, s));
        b += k;
        if (processed) *processed = b * B;
    }
    return TANX_OK;
}

// One block of the host-memory path: inputs at in[ch] + off, outputs at out[ch] + off.
static int32_t process_host_block(tanx_conv *cv, const float *const *in, float *const *out, size_t off,
                                  size_t n, const uint32_t *flags)
{
    const size_t B = cv->B;
    for (int ch = 0; ch < cv->C; ch++) std::memcpy(cv->hin + ch * B, in[ch] + off, B * sizeof(float));
    hipStream_t s = cv->ctx->proc;
    TANX_CHECK(hipMemcpyAsync(cv->din, cv->hin, cv->C * B * sizeof(float), hipMemcpyHostToDevice, s));
    int32_t rc = process_block(cv, cv->din, (long long)B, cv->dout, (long long)B, n, flags, false);
    if (rc != TANX_OK) return rc;
    TANX_CHECK(hipMemcpyAsync(cv->hout, cv->dout, cv->C * B * sizeof(float), hipMemcpyDeviceToHost, s));
    TANX_CHECK(hipEventRecord(cv->ev_out, s));
    if (!cv->fmode) {
        rc = enqueue_tail(cv);
        if (rc != TANX_OK) return rc;
    }
    TANX_CHECK(hipEventSynchronize(cv->ev_out));
    for (int ch = 0; ch < cv->C; ch++)
        if (cv->run_now[ch]) std::memcpy(out[ch] + off, cv->hout + ch * B, B * sizeof(float));
    return TANX_OK;
}

// Processes every whole block of n samples, one Process step per block
// (TrueAudioNext.h:208-210: the processed count is n rounded down to a
// multiple of the block size).
extern "C" int32_t tanx_conv_process(tanx_conv *cv, const float *const *in, float *const *out,
                                     size_t n, const uint32_t *flags, size_t *processed)
{
    if (!cv) return TANX_NOT_INITIALIZED;
    if (!in || !out) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> lk(cv->mtx);
    if (processed) *processed = 0;
    const size_t B = cv->B, nb = n / B;
    if (nb == 0) return TANX_OK;
    for (int ch = 0; ch < cv->C; ch++)
        if (!in[ch] || !out[ch]) return TANX_INVALID_ARG;
    DeviceGuard g(cv->ctx->device);
    for (size_t b = 0; b < nb; b++) {
        int32_t rc = process_host_block(cv, in, out, b * B, B, flags);
        if (rc != TANX_OK) return rc;
        if (processed) *processed = (b + 1) * B;
    }
    return TANX_OK;
}

extern "C" int32_t tanx_conv_process_device(tanx_conv *cv, const float *d_in, size_t in_stride,
                                            float *d_out, size_t out_stride, size_t n,
                                            const uint32_t *flags, size_t *processed)
{
    if (!cv) return TANX_NOT_INITIALIZED;
    if (!d_in || !d_out) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> lk(cv->mtx);
    if (processed) *processed = 0;
    const size_t B = cv->B, nb = n / B;
    if (nb == 0) return TANX_OK;
    if (in_stride < nb * B || out_stride < nb * B) return TANX_INVALID_ARG;
    DeviceGuard g(cv->ctx->device);
    if (cv->use_graph && cv->fused && !cv->fmode && !cv->timing && !cv->nu && !cv->td && cv->kp == 1 && !flags &&
        nb >= 2)
        return process_device_graph(cv, d_in, in_stride, d_out, out_stride, nb, processed);
    HeadQueue hq;
    const bool fuse = cv->fused && !cv->fmode;
    for (size_t b = 0; b < nb; b++) {
        int32_t rc = (fuse && cv->kp > 1)
                         ? process_block_kp(cv, d_in + b * B, (long long)in_stride, d_out + b * B,
                                            (long long)out_stride, B, flags, hq)
                         : process_block(cv, d_in + b * B, (long long)in_stride, d_out + b * B,
                                         (long long)out_stride, B, flags, fuse);
        if (rc != TANX_OK) return rc;
        if (!cv->fmode && !fuse) {
            rc = enqueue_tail(cv);
            if (rc != TANX_OK) return rc;
        }
        if (processed) *processed = (b + 1) * B;
    }
    return flush_heads(cv, hq);
}

/* ProcessDirect (ConvolutionImpl.cpp:719-762): time-domain objects only.
 * The IR comes with the call (channel c: ir[c], taps [first, last) of
 * nz_first_last[2c], [2c+1], or [0, len) when it is NULL); min(n, len)
 * samples are convolved against the channel's history, in chunks of at most
 * B samples, and the history advances; every channel is processed. */
extern "C" int32_t tanx_conv_process_direct(tanx_conv *cv, const float *const *ir, const float *const *in,
                                            float *const *out, size_t n, size_t *processed,
                                            const int32_t *nz_first_last)
{
    if (!cv) return TANX_NOT_INITIALIZED;
    if (processed) *processed = 0;
    if (!cv->td) return TANX_NOT_SUPPORTED;
    if (!ir || !in || !out) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> lk(cv->mtx);
    DeviceGuard g(cv->ctx->device);
    const size_t C = cv->C, len = cv->len, B = cv->B;
    n = std::min(n, len);
    for (size_t c = 0; c < C; c++)
        if (!ir[c] || !in[c] || !out[c]) return TANX_INVALID_ARG;
    hipStream_t s = cv->ctx->proc;
    // staging for the IRs (the host-update buffers), the ranges and an
    // all-running mask
    if (!cv->dir || !cv->hir) {
        const size_t bytes = C * len * sizeof(float);
        float *d = nullptr, *h = nullptr;
        hipError_t e1 = hipMalloc(&d, bytes);
        hipError_t e2 = e1 == hipSuccess ? hipHostMalloc(&h, bytes) : e1;
        if (e2 != hipSuccess) {
            if (d) (void)hipFree(d);
            return map_hip(e2);
        }
        cv->dir = d;
        cv->hir = h;
        cv->dbytes += bytes;
    }
    if (!cv->d_nz) {
        TANX_CHECK(hipMalloc(&cv->d_nz, C * 2 * sizeof(int) + C));
        cv->dbytes += C * 2 * sizeof(int) + C;
    }
    // the previous direct call's staging may still be read
    TANX_CHECK(hipStreamSynchronize(s));
    std::vector<int> nz(2 * C);
    for (size_t c = 0; c < C; c++) {
        int f0 = 0, f1 = (int)len;
        if (nz_first_last) {
            f0 = std::max(0, std::min((int)len, (int)nz_first_last[2 * c]));
            f1 = std::max(f0, std::min((int)len, (int)nz_first_last[2 * c + 1]));
        }
        nz[2 * c] = f0;
        nz[2 * c + 1] = f1;
        std::memcpy(cv->hir + c * len + f0, ir[c] + f0, (size_t)(f1 - f0) * sizeof(float));
        TANX_CHECK(hipMemcpyAsync(cv->dir + c * len + f0, cv->hir + c * len + f0, (size_t)(f1 - f0) * sizeof(float),
                                  hipMemcpyHostToDevice, s));
    }
    std::vector<uint8_t> ones(C, 1);
    uint8_t *d_ones = reinterpret_cast<uint8_t *>(cv->d_nz + 2 * C);
    TANX_CHECK(hipMemcpyAsync(cv->d_nz, nz.data(), 2 * C * sizeof(int), hipMemcpyHostToDevice, s));
    TANX_CHECK(hipMemcpyAsync(d_ones, ones.data(), C, hipMemcpyHostToDevice, s));
    for (size_t off = 0; off < n; off += B) {
        const size_t m = std::min(B, n - off);
        for (size_t c = 0; c < C; c++) std::memcpy(cv->hin + c * B, in[c] + off, m * sizeof(float));
        TANX_CHECK(hipMemcpyAsync(cv->din, cv->hin, C * B * sizeof(float), hipMemcpyHostToDevice, s));
        TdArgs t;
        td_common(cv, t);
        t.run = d_ones;
        t.n = (int)m;
        t.xin = cv->din;
        t.in_stride = (long long)B;
        t.yout = cv->dout;
        t.out_stride = (long long)B;
        t.pos = cv->tdpos;
        t.Ha = cv->dir;
        t.nz = cv->d_nz;
        TANX_CHECK(launch_td(t, s));
        TANX_CHECK(hipMemcpyAsync(cv->hout, cv->dout, C * B * sizeof(float), hipMemcpyDeviceToHost, s));
        TANX_CHECK(hipStreamSynchronize(s));
        for (size_t c = 0; c < C; c++) std::memcpy(out[c] + off, cv->hout + c * B, m * sizeof(float));
        cv->tdpos = (int)((cv->tdpos + m) & (len - 1));
    }
    if (processed) *processed = n;
    return TANX_OK;
}

extern "C" int32_t tanx_conv_process_direct_device(tanx_conv *cv, const float *ir, size_t ir_stride, const float *in,
                                                   size_t in_stride, float *out, size_t out_stride, size_t n,
                                                   size_t *processed, const int32_t *nz_first_last)
{
    if (!cv) return TANX_NOT_INITIALIZED;
    if (processed) *processed = 0;
    if (!cv->td) return TANX_NOT_SUPPORTED;
    if (!ir || !in || !out) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> lk(cv->mtx);
    DeviceGuard g(cv->ctx->device);
    const size_t C = cv->C, len = cv->len, B = cv->B;
    n = std::min(n, len);
    if (C > 1 && out_stride < n) return TANX_INVALID_ARG;  // channels' outputs must not overlap
    if (!cv->d_nz) {
        TANX_CHECK(hipMalloc(&cv->d_nz, C * 2 * sizeof(int) + C));
        cv->dbytes += C * 2 * sizeof(int) + C;
    }
    hipStream_t s = cv->ctx->proc;
    std::vector<int> nz(2 * C);
    for (size_t c = 0; c < C; c++) {
        int f0 = 0, f1 = (int)len;
        if (nz_first_last) {
            f0 = std::max(0, std::min((int)len, (int)nz_first_last[2 * c]));
            f1 = std::max(f0, std::min((int)len, (int)nz_first_last[2 * c + 1]));
        }
        nz[2 * c] = f0;
        nz[2 * c + 1] = f1;
    }
    // stream-ordered after the previous call's kernels read the ranges
    std::vector<uint8_t> ones(C, 1);
    uint8_t *d_ones = reinterpret_cast<uint8_t *>(cv->d_nz + 2 * C);
    TANX_CHECK(hipMemcpyAsync(cv->d_nz, nz.data(), 2 * C * sizeof(int), hipMemcpyHostToDevice, s));
    TANX_CHECK(hipMemcpyAsync(d_ones, ones.data(), C, hipMemcpyHostToDevice, s));
    for (size_t off = 0; off < n; off += B) {
        const size_t m = std::min(B, n - off);
        TdArgs t;
        td_common(cv, t);
        t.run = d_ones;
        t.n = (int)m;
        t.xin = in + off;
        t.in_stride = (long long)in_stride;
        t.yout = out + off;
        t.out_stride = (long long)out_stride;
        t.pos = cv->tdpos;
        t.Ha = ir;
        t.h_stride = (long long)ir_stride;
        t.nz = cv->d_nz;
        TANX_CHECK(launch_td(t, s));
        cv->tdpos = (int)((cv->tdpos + m) & (len - 1));
    }
    if (processed) *processed = n;
    return TANX_OK;
}

extern "C" int32_t tanx_conv_finalize(tanx_conv *cv)
{
    if (!cv) return TANX_NOT_INITIALIZED;
    std::lock_guard<std::mutex> lk(cv->mtx);
    DeviceGuard g(cv->ctx->device);
    return enqueue_tail(cv);
}

// Stages an IR update.  Slot choice: while a crossfade is armed (delayed),
// the reference blocks the caller until the crossfade block has run
// (ConvolutionImpl.cpp:425-428); here the update is written into the slot
// that becomes the update slot after that crossfade, and handed off by the
// first regular block after it, which is the order the reference produces
// for a caller on another thread.
// An IR transform of one object's update: partition size B, nchan listed
// channels (a grouped update launches many in one kernel).
struct IrJob {
    IrArgs a;
    int B, nchan;
};

// defer: where the update's IR transforms go instead of being launched
// (the caller launches them before anything else on the process stream).
static int32_t update_common(tanx_conv *cv, const float *const *hir, const float *dir, size_t dstride,
                             size_t n, const uint32_t *flags, uint32_t opflags,
                             std::vector<IrJob> *defer = nullptr)
{
    // the process stream: blocks enqueued before the update finish reading
    // the target slot before it is written, and blocks after it see the
    // new spectra, with no cross-stream wait (a wait costs the process
    // stream an idle gap of several microseconds)
    tanx_ctx *ctx = cv->ctx;
    hipStream_t s = ctx->proc;
    if (n > (size_t)cv->L) return TANX_INVALID_ARG;
    if (cv->stage_busy) {
        TANX_CHECK(hipEventSynchronize(cv->ev_upd));  // staging buffers free
        cv->stage_busy = false;
    }
    for (int ch = 0; ch < cv->C; ch++) {
        if (flags && (flags[ch] & TANX_CHANNEL_FLAG_FLUSH_STREAM)) {
            int32_t rc = flush_channel(cv, ch);
            if (rc != TANX_OK) return rc;
        }
    }
    // an armed switch holds the current and the armed filter: with two
    // slots there is none left for this update
    if (cv->delayed && cv->nslots == 2) return TANX_INPUT_FULL;
    const int target = cv->delayed ? cv->prev : cv->upd;
    const int keep = cv->delayed ? cv->upd : cv->cur;
    int nupd = 0;
    std::vector<int> keep_list;
    for (int ch = 0; ch < cv->C; ch++) {
        if (flags && (flags[ch] & TANX_CHANNEL_FLAG_STOP_INPUT)) {
            if (!cv->staged[ch]) keep_list.push_back(ch);
            continue;
        }
        cv->hchan[nupd++] = ch;
    }
    const size_t chanBytes = (size_t)cv->P * cv->row_bytes();
    bool staged = hir != nullptr && nupd > 0;  // pinned staging copies enqueued
    const int *chan = nullptr;
    if (nupd > 0) {
        const float *src = dir;
        long long stride = (long long)dstride;
        if (hir) {
            if (!cv->dir || !cv->hir) {
                const size_t bytes = (size_t)cv->C * cv->len * sizeof(float);
                float *d = nullptr, *h = nullptr;
                hipError_t e1 = hipMalloc(&d, bytes);
                hipError_t e2 = e1 == hipSuccess ? hipHostMalloc(&h, bytes) : e1;
                if (e2 != hipSuccess) {
                    if (d) (void)hipFree(d);
                    return map_hip(e2);
                }
                cv->dir = d;
                cv->hir = h;
                cv->dbytes += bytes;
            }
            for (int k = 0; k < nupd; k++) {
                const int ch = cv->hchan[k];
                if (!hir[ch]) return TANX_INVALID_ARG;
                std::memcpy(cv->hir + (size_t)ch * cv->len, hir[ch], n * sizeof(float));
                TANX_CHECK(hipMemcpyAsync(cv->dir + (size_t)ch * cv->len, cv->hir + (size_t)ch * cv->len,
                                          n * sizeof(float), hipMemcpyHostToDevice, s));
            }
            src = cv->dir;
            stride = cv->len;
        }
        // every channel: the kernels take row y as channel y, no list is copied
        if (nupd < cv->C) {
            TANX_CHECK(hipMemcpyAsync(cv->dchan, cv->hchan, nupd * sizeof(int), hipMemcpyHostToDevice, s));
            staged = true;
        }
        chan = nupd < cv->C ? cv->dchan : nullptr;
        if (cv->td) {
            TANX_CHECK(launch_td_load(src, stride, (long long)n, chan, nupd, cv->td_slot(target), cv->len,
                                      s));
            for (int k = 0; k < nupd; k++) cv->staged[cv->hchan[k]] = 1;
        }
    }
    if (nupd > 0 && !cv->td) {
        const float *src = dir;
        long long stride = (long long)dstride;
        if (hir) {
            src = cv->dir;
            stride = cv->len;
        }
        IrArgs a{};
        a.ir = src;
        a.stride = stride;
        a.n = (long long)n;
        a.chan = chan;
        a.H = cv->slot_ptr(target);
        a.S = cv->scale_ptr(target);
        a.irtype = cv->irtype;
        a.P = cv->P;
        const int lB = __builtin_ctz((unsigned)cv->B);
        a.twM = tw_table(ctx->tw, lB);
        a.twN = tw_table(ctx->tw, lB + 1);
        if (defer) defer->push_back(IrJob{a, cv->B, nupd});
        else TANX_CHECK(launch_ir_transform(cv->B, nupd, a, s));
        if (cv->nu) {
            // taps [N, len): frame partitions of N
            IrArgs b = a;
            b.ir = src + cv->N;
            b.n = (long long)n > cv->N ? (long long)n - cv->N : 0;
            b.H = cv->bslot_ptr(target);
            b.S = cv->bscale_ptr(target);
            b.P = cv->PB;
            const int lN = __builtin_ctz((unsigned)cv->N);
            b.twM = tw_table(ctx->tw, lN);
            b.twN = tw_table(ctx->tw, lN + 1);
            if (defer) defer->push_back(IrJob{b, cv->N, nupd});
            else TANX_CHECK(launch_ir_transform(cv->N, nupd, b, s));
        }
        for (int k = 0; k < nupd; k++) cv->staged[cv->hchan[k]] = 1;
    }
    // channels not updated keep the IR that is current when the update is
    // handed off (ConvolutionImpl.cpp:2149-2158)
    for (int ch : keep_list) {
        if (cv->td) {
            TANX_CHECK(hipMemcpyAsync(cv->td_slot(target) + (size_t)ch * cv->len, cv->td_slot(keep) + (size_t)ch * cv->len,
                                      (size_t)cv->len * sizeof(float), hipMemcpyDeviceToDevice, s));
            continue;
        }
        TANX_CHECK(hipMemcpyAsync(static_cast<char *>(cv->slot_ptr(target)) + (size_t)ch * chanBytes,
                                  static_cast<char *>(cv->slot_ptr(keep)) + (size_t)ch * chanBytes, chanBytes,
                                  hipMemcpyDeviceToDevice, s));
        TANX_CHECK(hipMemcpyAsync(cv->scale_ptr(target) + (size_t)ch * cv->P, cv->scale_ptr(keep) + (size_t)ch * cv->P,
                                  (size_t)cv->P * sizeof(float), hipMemcpyDeviceToDevice, s));
        if (cv->nu) {
            const size_t cb = (size_t)cv->PB * cv->brow_bytes();
            TANX_CHECK(hipMemcpyAsync(static_cast<char *>(cv->bslot_ptr(target)) + (size_t)ch * cb,
                                      static_cast<char *>(cv->bslot_ptr(keep)) + (size_t)ch * cb, cb,
                                      hipMemcpyDeviceToDevice, s));
            TANX_CHECK(hipMemcpyAsync(cv->bscale_ptr(target) + (size_t)ch * cv->PB,
                                      cv->bscale_ptr(keep) + (size_t)ch * cv->PB, (size_t)cv->PB * sizeof(float),
                                      hipMemcpyDeviceToDevice, s));
        }
    }
    const bool block = (opflags & TANX_OPERATION_FLAG_BLOCK_UNTIL_READY) != 0;
    if (staged || block) TANX_CHECK(hipEventRecord(cv->ev_upd, s));
    cv->stage_busy = staged;
    cv->pending = true;
    if (block) {
        TANX_CHECK(hipEventSynchronize(cv->ev_upd));
        cv->stage_busy = false;
    }
    return TANX_OK;
}

// ---------------------------------------------------------------------------
// group launches
// ---------------------------------------------------------------------------
static int32_t group_staging(tanx_ctx *ctx, size_t bytes, char **host, int *slot)
{
    if (!ctx->g_ev[0]) {
        for (auto &e : ctx->g_ev) {
            TANX_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            TANX_CHECK(hipEventRecord(e, ctx->proc));
        }
    }
    if (ctx->g_cap < bytes) {
        // queued group kernels may still read the device copy
        TANX_CHECK(hipStreamSynchronize(ctx->proc));
        if (ctx->g_host) (void)hipHostFree(ctx->g_host);
        if (ctx->g_dev) (void)hipFree(ctx->g_dev);
        ctx->g_host = nullptr;
        ctx->g_dev = nullptr;
        ctx->g_cap = 0;
        size_t cap = 4096;
        while (cap < bytes) cap *= 2;
        TANX_CHECK(hipHostMalloc(&ctx->g_host, cap * 4));
        TANX_CHECK(hipMalloc(&ctx->g_dev, cap));
        ctx->g_cap = cap;
    }
    *slot = ctx->g_ring;
    ctx->g_ring = (ctx->g_ring + 1) % 4;
    TANX_CHECK(hipEventSynchronize(ctx->g_ev[*slot]));
    *host = ctx->g_host + (size_t)(*slot) * ctx->g_cap;
    return TANX_OK;
}

// Checks the arguments of a group call; returns the context and IR type.
static int32_t group_check(uint32_t n, tanx_conv *const *convs, const float *const *d_in, const size_t *in_stride,
                           float *const *d_out, const size_t *out_stride, tanx_ctx **ctx_out, int *irtype_out)
{
    if (!convs || !d_in || !d_out || !in_stride || !out_stride) return TANX_INVALID_ARG;
    tanx_ctx *ctx = convs[0] ? convs[0]->ctx : nullptr;
    if (!ctx) return TANX_NOT_INITIALIZED;
    const int irtype = convs[0]->irtype;
    for (uint32_t i = 0; i < n; i++) {
        tanx_conv *cv = convs[i];
        if (!cv) return TANX_NOT_INITIALIZED;
        if (cv->ctx != ctx || cv->irtype != irtype || cv->B > kGroupMaxB || cv->fmode || cv->td)
            return TANX_NOT_SUPPORTED;
        if (!d_in[i] || !d_out[i] || in_stride[i] < (size_t)cv->B || out_stride[i] < (size_t)cv->B)
            return TANX_INVALID_ARG;
    }
    *ctx_out = ctx;
    *irtype_out = irtype;
    return TANX_OK;
}

// Where one group call's descriptors go: host staging and the device copy,
// both laid out [TickArgs x cap][int x cap][FrameArgs x cap][int x cap]
// (16-byte aligned pieces) for up to `cap` objects.
struct GroupStage {
    char *host;
    char *dev;
    size_t desc_bytes, start_bytes, fdesc_bytes, fstart_bytes;
    size_t total() const { return desc_bytes + start_bytes + fdesc_bytes + fstart_bytes; }
};

static GroupStage group_layout(size_t cap)
{
    GroupStage g{};
    g.desc_bytes = align16(sizeof(TickArgs) * cap);
    g.start_bytes = align16(sizeof(int) * cap);
    g.fdesc_bytes = align16(sizeof(FrameArgs) * cap);
    g.fstart_bytes = align16(sizeof(int) * cap);
    return g;
}

// Advances every object (locked by the caller) one block, in call order, and
// writes its launch descriptors into g.host, grouped by block size (the
// objects of one block size contiguous, in call order); then copies them to
// g.dev on the process stream and launches one group kernel per block size
// present and the frame kernels.  Objects already advanced are launched even
// if a later one fails (its error is returned).  staged: the event to record
// after the copy (the host staging is free again once it completes), or
// nullptr.  may_queue: the launches may take the context's work-queue form.
static int32_t group_run(tanx_ctx *ctx, int irtype, uint32_t n, tanx_conv *const *convs, const float *const *d_in,
                         const size_t *in_stride, float *const *d_out, const size_t *out_stride,
                         const uint32_t *const *flags, const GroupStage &g, hipEvent_t staged, bool may_queue)
{
    TickArgs *descs = reinterpret_cast<TickArgs *>(g.host);
    int *starts = reinterpret_cast<int *>(g.host + g.desc_bytes);
    FrameArgs *fdescs = reinterpret_cast<FrameArgs *>(g.host + g.desc_bytes + g.start_bytes);
    int *fstarts = reinterpret_cast<int *>(g.host + g.desc_bytes + g.start_bytes + g.fdesc_bytes);
    // block-size classes: descriptor range [first, first + placed) of each
    constexpr int kClasses = 4;
    struct Class {
        uint32_t first = 0, size = 0, placed = 0;
        int total = 0;
    } cls[kClasses];
    auto class_of = [](int B) { return B == 64 ? 0 : B == 128 ? 1 : B == 256 ? 2 : 3; };
    for (uint32_t i = 0; i < n; i++) cls[class_of(convs[i]->B)].size++;
    for (int c = 1; c < kClasses; c++) cls[c].first = cls[c - 1].first + cls[c - 1].size;
    uint32_t done = 0;
    int nframe = 0, ftotal = 0;
    int32_t rc = TANX_OK;
    uint64_t tick_bytes = 0, frame_bytes = 0;
    for (; done < n; done++) {
        tanx_conv *cv = convs[done];
        Class &k = cls[class_of(cv->B)];
        const uint32_t at = k.first + k.placed;
        int nf = 1;
        bool xf = false, fdue = false;
        rc = prepare_block(cv, d_in[done], (long long)in_stride[done], d_out[done], (long long)out_stride[done],
                           cv->B, flags ? flags[done] : nullptr, true, descs[at], &nf, &xf, &fdescs[nframe], &fdue);
        if (rc != TANX_OK) break;
        const TickArgs &a = descs[at];
        if (ctx->gtiming) {
            tick_bytes += head_bytes(cv, xf ? 2 : 1) + near_bytes(cv, a) + frame_mac_bytes(cv, a.bHb ? 2 : 1) +
                          (a.nchunk > 0 ? tail_bytes(cv, nf) : 0);
            if (fdue) frame_bytes += frame_kernel_bytes(cv, fdescs[nframe].Hb ? 2 : 1);
        }
        // a group streams every member's rows (group kernels ignore the flag)
        descs[at].cached = 0;
        tail_done(cv);
        k.placed++;
        if (fdue) {
            fstarts[nframe++] = ftotal;
            ftotal += cv->C;
        }
    }
    if (done > 0) {
        // each class's range splits by launch kind (role set, then pass
        // length) in place, each sub-range one launch with block offsets of
        // its own: lean 4, 8, 12, 16, crossfade 4, 8, 12, 16, general, heads
        // (objects without a pass this tick, whatever their pass length: a
        // launch's register count is paid by all of its workgroups)
        constexpr int kKinds = 10;
        auto kind_of = [](const TickArgs &a) {
            const int r = group_roles(a);
            if (r == kRolesHeads) return 9;
            if (r == kRolesGeneral) return 8;
            return 4 * (r == kRolesXfade) + pass_km(a.kp) / 4 - 1;
        };
        auto kind_roles = [](int kind) {
            return kind == 9 ? kRolesHeads : kind == 8 ? kRolesGeneral : kind >= 4 ? kRolesXfade : kRolesLean;
        };
        struct Launch {
            uint32_t first, count;
            int B, total, roles, km;
        } launches[kKinds * kClasses];
        int nl = 0;
        for (int c = 0; c < kClasses; c++) {
            const Class &k = cls[c];
            uint32_t lo = k.first;
            const uint32_t end = k.first + k.placed;
            for (int kind = 0; kind < kKinds && lo < end; kind++) {
                // move this kind's descriptors to the front of [lo, end)
                uint32_t mid = lo;
                for (uint32_t i = lo; i < end; i++)
                    if (kind_of(descs[i]) == kind) std::swap(descs[i], descs[mid++]);
                if (mid == lo) continue;
                Launch &l = launches[nl++];
                l = Launch{lo, mid - lo, 64 << c, 0, kind_roles(kind), kind < 8 ? 4 * (kind % 4 + 1) : 4};
                for (uint32_t i = lo; i < mid; i++) {
                    starts[i] = l.total;
                    l.total += group_blocks(descs[i], l.roles);
                }
                lo = mid;
            }
        }
        // the staging pieces are copied in one transfer at their offsets
        TANX_CHECK(hipMemcpyAsync(g.dev, g.host, g.total(), hipMemcpyHostToDevice, ctx->proc));
        if (staged) TANX_CHECK(hipEventRecord(staged, ctx->proc));
        hipEvent_t t0 = nullptr, t1 = nullptr;
        if (ctx->gtiming) {
            TANX_CHECK(ctx->gt_tick.acquire(&t0, &t1));
        }
        const bool queue = may_queue && ctx->gqueue;
        if (queue) {
            static_assert(kKinds * kClasses * sizeof(int) % 16 == 0, "counters in whole float4");
            if (!ctx->q_next) TANX_CHECK(hipMalloc(&ctx->q_next, kKinds * kClasses * sizeof(int)));
            TANX_CHECK(launch_zero(ctx->q_next, align16((size_t)nl * sizeof(int)), ctx->proc));
        }
        for (int i = 0; i < nl; i++) {
            const Launch &l = launches[i];
            const TickArgs *dd = reinterpret_cast<const TickArgs *>(g.dev) + l.first;
            const int *ds = reinterpret_cast<const int *>(g.dev + g.desc_bytes) + l.first;
            // the first launch stamps the start event, the last one the end
            if (queue)
                TANX_CHECK(launch_group_queue(irtype, l.B, l.roles, l.km, dd, ds, (int)l.count, l.total,
                                              ctx->q_next + i, ctx->proc, i == 0 ? t0 : nullptr,
                                              i == nl - 1 ? t1 : nullptr));
            else
                TANX_CHECK(launch_group_tick(irtype, l.B, l.roles, l.km, dd, ds, (int)l.count, (unsigned)l.total,
                                             ctx->proc, i == 0 ? t0 : nullptr, i == nl - 1 ? t1 : nullptr));
        }
        if (ctx->gtiming) {
            ctx->gt_tick.commit();
            ctx->g_tick_bytes += tick_bytes;
            ctx->g_objects += done;
        }
        if (nframe > 0) {
            hipEvent_t f0 = nullptr, f1 = nullptr;
            if (ctx->gtiming) {
                TANX_CHECK(ctx->gt_frame.acquire(&f0, &f1));
            }
            TANX_CHECK(launch_group_frame(
                irtype, reinterpret_cast<const FrameArgs *>(g.dev + g.desc_bytes + g.start_bytes),
                reinterpret_cast<const int *>(g.dev + g.desc_bytes + g.start_bytes + g.fdesc_bytes), nframe,
                (unsigned)ftotal, ctx->proc, f0, f1));
            if (ctx->gtiming) {
                ctx->gt_frame.commit();
                ctx->g_frame_bytes += frame_bytes;
            }
        }
    } else if (staged) {
        TANX_CHECK(hipEventRecord(staged, ctx->proc));
    }
    return rc;
}

extern "C" int32_t tanx_group_process_device(uint32_t n, tanx_conv *const *convs, const float *const *d_in,
                                             const size_t *in_stride, float *const *d_out,
                                             const size_t *out_stride, const uint32_t *const *flags)
{
    if (n == 0) return TANX_OK;
    tanx_ctx *ctx = nullptr;
    int irtype = 0;
    int32_t rc = group_check(n, convs, d_in, in_stride, d_out, out_stride, &ctx, &irtype);
    if (rc != TANX_OK) return rc;
    // lock every object (address order) and the context's group staging
    std::vector<tanx_conv *> order(convs, convs + n);
    std::sort(order.begin(), order.end());
    if (std::adjacent_find(order.begin(), order.end()) != order.end()) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> glk(ctx->group_mtx);
    std::vector<std::unique_lock<std::mutex>> locks;
    locks.reserve(n);
    for (tanx_conv *cv : order) locks.emplace_back(cv->mtx);
    DeviceGuard g(ctx->device);
    GroupStage st = group_layout(n);
    int slot = 0;
    rc = group_staging(ctx, st.total(), &st.host, &slot);
    if (rc != TANX_OK) return rc;
    st.dev = ctx->g_dev;
    return group_run(ctx, irtype, n, convs, d_in, in_stride, d_out, out_stride, flags, st, ctx->g_ev[slot], true);
}

extern "C" int32_t tanx_group_update_td_device(uint32_t n, tanx_conv *const *convs, const float *const *d_irs,
                                               const size_t *ir_stride, const size_t *ir_len,
                                               const uint32_t *const *flags, uint32_t opflags)
{
    if (n == 0) return TANX_OK;
    if (!convs || !d_irs || !ir_stride || !ir_len) return TANX_INVALID_ARG;
    tanx_ctx *ctx = convs[0] ? convs[0]->ctx : nullptr;
    if (!ctx) return TANX_NOT_INITIALIZED;
    for (uint32_t i = 0; i < n; i++) {
        if (!convs[i]) return TANX_NOT_INITIALIZED;
        if (convs[i]->ctx != ctx || convs[i]->td) return TANX_NOT_SUPPORTED;
        if (!d_irs[i] || ir_stride[i] < ir_len[i]) return TANX_INVALID_ARG;
    }
    std::vector<tanx_conv *> order(convs, convs + n);
    std::sort(order.begin(), order.end());
    if (std::adjacent_find(order.begin(), order.end()) != order.end()) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> glk(ctx->group_mtx);
    std::vector<std::unique_lock<std::mutex>> locks;
    locks.reserve(n);
    for (tanx_conv *cv : order) locks.emplace_back(cv->mtx);
    DeviceGuard g(ctx->device);
    // each object's bookkeeping, slot choice and copies, in call order; its
    // transforms are collected (a failure leaves the earlier objects'
    // transforms to be launched below, as the group process call does)
    const uint32_t opf = opflags & ~uint32_t(TANX_OPERATION_FLAG_BLOCK_UNTIL_READY);
    std::vector<IrJob> jobs;
    int32_t rc = TANX_OK;
    for (uint32_t i = 0; i < n && rc == TANX_OK; i++)
        rc = update_common(convs[i], nullptr, d_irs[i], ir_stride[i], ir_len[i], flags ? flags[i] : nullptr, opf,
                           &jobs);
    if (!jobs.empty()) {
        // one launch per (partition size, IR type): descriptors and block
        // offsets staged in one copy
        std::stable_sort(jobs.begin(), jobs.end(), [](const IrJob &x, const IrJob &y) {
            return x.B != y.B ? x.B < y.B : x.a.irtype < y.a.irtype;
        });
        const size_t nj = jobs.size(), db = align16(sizeof(IrArgs) * nj), sb = align16(sizeof(int) * nj);
        char *host = nullptr;
        int slot = 0;
        const int32_t rs = group_staging(ctx, db + sb, &host, &slot);
        if (rs != TANX_OK) return rs;
        IrArgs *descs = reinterpret_cast<IrArgs *>(host);
        int *starts = reinterpret_cast<int *>(host + db);
        struct Launch {
            size_t first, count;
            int B, irtype;
            long long blocks;
        };
        std::vector<Launch> launches;
        for (size_t j = 0; j < nj; j++) {
            const IrJob &jb = jobs[j];
            if (launches.empty() || launches.back().B != jb.B || launches.back().irtype != jb.a.irtype)
                launches.push_back(Launch{j, 0, jb.B, jb.a.irtype, 0});
            Launch &l = launches.back();
            descs[j] = jb.a;
            starts[j] = (int)l.blocks;
            l.blocks += ir_transform_blocks(jb.B, jb.a.irtype, jb.a.P, jb.nchan);
            l.count++;
        }
        for (const Launch &l : launches)
            if (l.blocks > INT32_MAX) return TANX_INVALID_ARG;
        TANX_CHECK(hipMemcpyAsync(ctx->g_dev, host, db + sb, hipMemcpyHostToDevice, ctx->proc));
        TANX_CHECK(hipEventRecord(ctx->g_ev[slot], ctx->proc));
        for (const Launch &l : launches)
            TANX_CHECK(launch_ir_transform_group(l.B, l.irtype, reinterpret_cast<const IrArgs *>(ctx->g_dev) + l.first,
                                                 reinterpret_cast<const int *>(ctx->g_dev + db) + l.first,
                                                 (int)l.count, (unsigned)l.blocks, ctx->proc));
    }
    if (rc == TANX_OK && (opflags & TANX_OPERATION_FLAG_BLOCK_UNTIL_READY))
        TANX_CHECK(hipStreamSynchronize(ctx->proc));
    return rc;
}

extern "C" int32_t tanx_ctx_enable_group_timing(tanx_ctx *ctx, uint32_t enable)
{
    if (!ctx) return TANX_INVALID_POINTER;
    std::lock_guard<std::mutex> glk(ctx->group_mtx);
    DeviceGuard g(ctx->device);
    if (enable && ctx->gt_tick.start.empty()) {
        if (!ctx->gt_tick.init() || !ctx->gt_frame.init()) return TANX_UNEXPECTED;
    }
    ctx->gtiming = enable != 0;
    return TANX_OK;
}

extern "C" int32_t tanx_ctx_set_group_queue(tanx_ctx *ctx, uint32_t enable)
{
    if (!ctx) return TANX_INVALID_POINTER;
    std::lock_guard<std::mutex> glk(ctx->group_mtx);
    ctx->gqueue = enable != 0;
    return TANX_OK;
}

extern "C" int32_t tanx_ctx_read_group_timing(tanx_ctx *ctx, tanx_kernel_timing *out, uint32_t reset)
{
    if (!ctx || !out) return TANX_INVALID_POINTER;
    std::lock_guard<std::mutex> glk(ctx->group_mtx);
    DeviceGuard g(ctx->device);
    TANX_CHECK(ctx->gt_tick.drain());
    TANX_CHECK(ctx->gt_frame.drain());
    out->head_ms = ctx->gt_tick.ms;
    out->tail_ms = ctx->gt_frame.ms;
    out->head_launches = ctx->gt_tick.count;
    out->tail_launches = ctx->gt_frame.count;
    out->head_alg_bytes = ctx->g_tick_bytes;
    out->tail_alg_bytes = ctx->g_frame_bytes;
    out->blocks = ctx->g_objects;
    if (reset) {
        ctx->gt_tick.ms = ctx->gt_frame.ms = 0.0;
        ctx->gt_tick.count = ctx->gt_frame.count = 0;
        ctx->g_tick_bytes = ctx->g_frame_bytes = ctx->g_objects = 0;
    }
    return TANX_OK;
}

// Host staging of the allocation-free variant (pageable: a copy from it has
// consumed it when hipMemcpyAsync returns), its lock order array and its
// object locks, all static: the entry point allocates nothing.
namespace {
std::mutex g_noalloc_mtx;
alignas(16) char g_noalloc_host[kNoAllocScratchBytes];
tanx_conv *g_noalloc_order[kNoAllocGroupMax];
}  // namespace

extern "C" int32_t tanx_group_process_device_noalloc(uint32_t n, tanx_conv *const *convs,
                                                     const float *const *d_in, const size_t *in_stride,
                                                     float *const *d_out, const size_t *out_stride,
                                                     const uint32_t *const *flags)
{
    if (n == 0) return TANX_OK;
    if (n > (uint32_t)kNoAllocGroupMax) return TANX_INVALID_ARG;  // beyond the scratch array
    tanx_ctx *ctx = nullptr;
    int irtype = 0;
    int32_t rc = group_check(n, convs, d_in, in_stride, d_out, out_stride, &ctx, &irtype);
    if (rc != TANX_OK) return rc;
    char *dev = static_cast<char *>(noalloc_scratch(ctx->device));
    if (!dev) return TANX_NOT_INITIALIZED;  // the scratch address was not resolved at context creation
    // one call at a time per process: the static staging, order and locks
    std::lock_guard<std::mutex> slk(g_noalloc_mtx);
    std::copy(convs, convs + n, g_noalloc_order);
    std::sort(g_noalloc_order, g_noalloc_order + n);
    if (std::adjacent_find(g_noalloc_order, g_noalloc_order + n) != g_noalloc_order + n) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> glk(ctx->group_mtx);
    std::unique_lock<std::mutex> locks[kNoAllocGroupMax];
    for (uint32_t i = 0; i < n; i++) locks[i] = std::unique_lock<std::mutex>(g_noalloc_order[i]->mtx);
    DeviceGuard g(ctx->device);
    // packed for n objects: within the scratch array, laid out for kNoAllocGroupMax
    GroupStage st = group_layout(n);
    st.host = g_noalloc_host;
    st.dev = dev;
    // the plain group launches: the work-queue form keeps its counters in an
    // allocation of the context
    return group_run(ctx, irtype, n, convs, d_in, in_stride, d_out, out_stride, flags, st, nullptr, false);
}

extern "C" int32_t tanx_conv_update_td(tanx_conv *cv, const float *const *ir, size_t n,
                                       const uint32_t *flags, uint32_t opflags)
{
    if (!cv) return TANX_NOT_INITIALIZED;
    if (!ir) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> lk(cv->mtx);
    DeviceGuard g(cv->ctx->device);
    return update_common(cv, ir, nullptr, 0, n, flags, opflags);
}

extern "C" int32_t tanx_conv_update_td_device(tanx_conv *cv, const float *d_ir, size_t stride, size_t n,
                                              const uint32_t *flags, uint32_t opflags)
{
    if (!cv) return TANX_NOT_INITIALIZED;
    if (!d_ir || stride < n) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> lk(cv->mtx);
    DeviceGuard g(cv->ctx->device);
    return update_common(cv, nullptr, d_ir, stride, n, flags, opflags);
}

extern "C" int32_t tanx_conv_flush(tanx_conv *cv, uint32_t channel)
{
    if (!cv) return TANX_NOT_INITIALIZED;
    if (channel >= (uint32_t)cv->C) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> lk(cv->mtx);
    DeviceGuard g(cv->ctx->device);
    return flush_channel(cv, (int)channel);
}

extern "C" int32_t tanx_conv_get_next_free_channel(tanx_conv *cv, uint32_t *channel, const uint32_t *flags)
{
    if (!cv) return TANX_NOT_INITIALIZED;
    if (!channel) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> lk(cv->mtx);
    *channel = (uint32_t)-1;
    if (!flags) return TANX_NOT_FOUND;
    for (int ch = 0; ch < cv->C; ch++) {
        if ((flags[ch] & TANX_CHANNEL_FLAG_STOP_INPUT) && cv->tail_left[ch] <= 0) {
            *channel = (uint32_t)ch;
            return TANX_OK;
        }
    }
    return TANX_NOT_FOUND;
}

extern "C" int32_t tanx_conv_synchronize(tanx_conv *cv)
{
    if (!cv) return TANX_NOT_INITIALIZED;
    DeviceGuard g(cv->ctx->device);
    TANX_CHECK(hipStreamSynchronize(cv->ctx->proc));
    TANX_CHECK(hipStreamSynchronize(cv->ctx->upd));
    return TANX_OK;
}

extern "C" int32_t tanx_conv_get_info(const tanx_conv *cv, tanx_conv_info *info)
{
    if (!cv || !info) return TANX_INVALID_POINTER;
    info->ir_len = (uint32_t)cv->len;
    info->block = (uint32_t)cv->B;
    info->channels = (uint32_t)cv->C;
    info->partitions = (uint32_t)cv->P;
    info->tail_chunks = (uint32_t)cv->nchunk;
    info->finalize_mode = cv->fmode ? 1u : 0u;
    info->ir_precision = cv->irtype == IR_F16 ? TANX_IR_FP16 : cv->irtype == IR_BF16 ? TANX_IR_BF16 : TANX_IR_FP32;
    info->device_bytes = cv->dbytes;
    info->ticks = cv->ticks;
    info->crossfades = cv->xfades;
    info->nu_multiple = (uint32_t)cv->M;
    info->pass_blocks = (uint32_t)cv->kp;
    info->cached_rows = cached_rows(cv) ? 1u : 0u;
    info->ir_slots = (uint32_t)cv->nslots;
    return TANX_OK;
}

extern "C" int32_t tanx_conv_set_tail_chunks(tanx_conv *cv, uint32_t chunks)
{
    if (!cv) return TANX_NOT_INITIALIZED;
    std::lock_guard<std::mutex> lk(cv->mtx);
    set_chunks(cv, chunks == 0 ? auto_chunks(cv->C, cv->P) : (int)chunks);
    return TANX_OK;
}

extern "C" int32_t tanx_conv_enable_timing(tanx_conv *cv, uint32_t enable)
{
    if (!cv) return TANX_NOT_INITIALIZED;
    std::lock_guard<std::mutex> lk(cv->mtx);
    DeviceGuard g(cv->ctx->device);
    if (enable && cv->t_head.start.empty()) {
        if (!cv->t_head.init() || !cv->t_tail.init()) return TANX_UNEXPECTED;
    }
    if (!enable && cv->timing) {
        TANX_CHECK(cv->t_head.drain());
        TANX_CHECK(cv->t_tail.drain());
    }
    cv->timing = enable != 0;
    return TANX_OK;
}

extern "C" int32_t tanx_conv_read_timing(tanx_conv *cv, tanx_kernel_timing *out, uint32_t reset)
{
    if (!cv || !out) return TANX_INVALID_POINTER;
    std::lock_guard<std::mutex> lk(cv->mtx);
    DeviceGuard g(cv->ctx->device);
    TANX_CHECK(cv->t_head.drain());
    TANX_CHECK(cv->t_tail.drain());
    out->head_ms = cv->t_head.ms;
    out->head_launches = cv->t_head.count;
    out->head_alg_bytes = cv->head_alg_bytes;
    out->tail_ms = cv->t_tail.ms;
    out->tail_launches = cv->t_tail.count;
    out->tail_alg_bytes = cv->tail_alg_bytes;
    out->blocks = cv->timed_blocks;
    if (reset) {
        cv->t_head.ms = cv->t_tail.ms = 0.0;
        cv->t_head.count = cv->t_tail.count = 0;
        cv->head_alg_bytes = cv->tail_alg_bytes = 0;
        cv->timed_blocks = 0;
    }
    return TANX_OK;
}

// Picks the tail chunking by timing dry-run fused launches.  Only on a fresh
// object (no block processed yet): the FDL, previous block and partial sums
// are all zero, the dry run feeds a zero block, and so every buffer it writes
// keeps its zero contents.
extern "C" int32_t tanx_conv_autotune(tanx_conv *cv, uint32_t *chosen, double *best_us)
{
    if (!cv) return TANX_NOT_INITIALIZED;
    std::lock_guard<std::mutex> lk(cv->mtx);
    if (cv->td) return TANX_NOT_SUPPORTED;  // no partitions to chunk
    if (cv->ticks != 0) return TANX_WRONG_STATE;
    DeviceGuard g(cv->ctx->device);
    hipStream_t s = cv->ctx->proc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TANX_CHECK(hipEventCreate(&e0));
    TANX_CHECK(hipEventCreate(&e1));
    const int keep = cv->nchunk;
    double best = 1e30;
    int best_n = keep;
    int32_t rc = TANX_OK;
    // zero input block (a host Process that failed may have staged data in din)
    rc = map_hip(launch_zero(cv->din, (size_t)cv->C * cv->B * sizeof(float), s));
    for (int n = 1; n <= cv->maxchunk && rc == TANX_OK; n *= 2) {
        set_chunks(cv, n);
        if (cv->nchunk != n) continue;
        // consecutive blocks as Process runs them (silent input): ring slots
        // and partial-sum buffers rotate, so every launch streams the whole
        // FDL and IR slot as a real block does
        auto block_args = [&](int t) {
            TickArgs a = tick_common(cv);
            a.nhead = cv->C;
            a.xin = cv->din;
            a.in_stride = cv->B;
            a.yout = cv->dout;
            a.out_stride = cv->B;
            a.xprev = cv->xprev;
            a.slot = t % cv->P;
            a.Ha = a.Hb = cv->slot_ptr(cv->cur);
            a.Sa = a.Sb = cv->scale_ptr(cv->cur);
            a.Ta = a.Tb = cv->tail_ptr(t & 1, 0);
            a.tchunk = cv->nchunk;
            a.xfade = 0;
            a.nchunk = cv->nchunk;
            a.pc = cv->pc;
            a.tbase = (t + 1) % cv->P;
            a.tHa = cv->slot_ptr(cv->cur);
            a.tHb = nullptr;
            a.tSa = cv->scale_ptr(cv->cur);
            a.tSb = nullptr;
            a.tTa = cv->tail_ptr((t + 1) & 1, 0);
            a.tTb = nullptr;
            if (cv->kp > 1) {
                // a pass every kp blocks (with that block's head), head-only launches between
                const int k = t % cv->kp;
                a.Ta = a.Tb = cv->tail_slot(t, 0);
                a.dA = k == 0 ? cv->kp : k;
                if (k == 0) {
                    a.kn = cv->kp;
                    a.kgap = 1;
                    for (int q = 0; q < kPassMax; q++) {
                        a.kTA[q] = q < cv->kp ? cv->tail_slot(t + 1 + q, 0) : nullptr;
                        a.kTB[q] = nullptr;
                    }
                } else {
                    a.nchunk = 0;
                }
            }
            return a;
        };
        const int reps = cv->C >= 4096 ? 4 : 32;
        for (int w = 0; w < 4 && rc == TANX_OK; w++) rc = map_hip(launch_tick(cv->B, 1, block_args(w), s));
        if (rc != TANX_OK) break;
        if ((rc = map_hip(hipEventRecord(e0, s))) != TANX_OK) break;
        for (int r = 0; r < reps && rc == TANX_OK; r++) rc = map_hip(launch_tick(cv->B, 1, block_args(4 + r), s));
        if (rc != TANX_OK) break;
        if ((rc = map_hip(hipEventRecord(e1, s))) != TANX_OK) break;
        if ((rc = map_hip(hipEventSynchronize(e1))) != TANX_OK) break;
        float ms = 0.f;
        if ((rc = map_hip(hipEventElapsedTime(&ms, e0, e1))) != TANX_OK) break;
        const double us = 1e3 * ms / reps;
        if (us < best) {
            best = us;
            best_n = n;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_chunks(cv, rc == TANX_OK ? best_n : keep);
    cv->tail_nchunk = cv->nchunk;
    if (chosen) *chosen = (uint32_t)cv->nchunk;
    if (best_us) *best_us = best;
    return rc;
}

extern "C" int32_t tanx_conv_set_graph(tanx_conv *cv, uint32_t use_graph)
{
    if (!cv) return TANX_NOT_INITIALIZED;
    std::lock_guard<std::mutex> lk(cv->mtx);
    cv->use_graph = use_graph != 0;
    return TANX_OK;
}

extern "C" int32_t tanx_conv_set_fused(tanx_conv *cv, uint32_t fused)
{
    if (!cv) return TANX_NOT_INITIALIZED;
    std::lock_guard<std::mutex> lk(cv->mtx);
    cv->fused = fused != 0;
    return TANX_OK;
}

extern "C" int32_t tanx_conv_last_was_crossfade(const tanx_conv *cv)
{
    return cv && cv->last_xfade ? 1 : 0;
}

extern "C" void *tanx_conv_process_stream(tanx_conv *cv)
{
    return cv ? (void *)cv->ctx->proc : nullptr;
}

// ---------------------------------------------------------------------------
// TANFFT
// ---------------------------------------------------------------------------
static int32_t fft_check(uint32_t dir, uint32_t log2len, uint32_t channels)
{
    if (channels == 0 || log2len == 0) return TANX_INVALID_ARG;
    if (dir > TANX_FFT_C2R_PLANAR_BACKWARD) return TANX_INVALID_ARG;
    if (log2len > (uint32_t)kFftLog2Max) return TANX_NOT_SUPPORTED;
    // planar layout ends at N + 9 floats of a 2N-float channel buffer
    if (dir >= TANX_FFT_R2C_PLANAR_FORWARD && log2len < 4) return TANX_INVALID_ARG;
    return TANX_OK;
}

// Launch arguments of a transform; sizes above one workgroup's LDS get the
// four-step scratch (grown as needed, kept by the context; the caller holds
// fft_mtx) and the tables of its column and row sizes.
static int32_t fft_args(tanx_ctx *ctx, uint32_t dir, uint32_t log2len, uint32_t channels, const float *in,
                        float *out, size_t stride, FftArgs &a)
{
    a = FftArgs{};
    a.in = in;
    a.out = out;
    a.stride = (long long)stride;
    a.mode = (int)dir;
    a.log2len = (int)log2len;
    const int lM = dir >= TANX_FFT_R2C_FORWARD ? (int)log2len - 1 : (int)log2len;
    if ((1 << lM) <= kFftLdsMax) {
        a.twM = tw_table(ctx->tw, lM);
        a.twN = tw_table(ctx->tw, lM + 1);
        return TANX_OK;
    }
    const size_t need = size_t(2) * channels * (size_t(1) << lM);
    if (ctx->fft_scratch_elems < need) {
        // the stream may still run a transform that uses the old scratch
        TANX_CHECK(hipStreamSynchronize(ctx->upd));
        if (ctx->fft_scratch) (void)hipFree(ctx->fft_scratch);
        ctx->fft_scratch = nullptr;
        ctx->fft_scratch_elems = 0;
        TANX_CHECK(hipMalloc(&ctx->fft_scratch, need * sizeof(float2)));
        ctx->fft_scratch_elems = need;
    }
    a.scratch = ctx->fft_scratch;
    const int l1 = lM / 2;
    a.tw1 = tw_table(ctx->tw, l1);
    a.tw2 = tw_table(ctx->tw, lM - l1);
    return TANX_OK;
}

extern "C" int32_t tanx_fft_transform_device(tanx_ctx *ctx, uint32_t direction, uint32_t log2len,
                                             uint32_t channels, const float *d_in, float *d_out, size_t stride)
{
    if (!ctx) return TANX_NOT_INITIALIZED;
    int32_t rc = fft_check(direction, log2len, channels);
    if (rc != TANX_OK) return rc;
    if (!d_in || !d_out || stride < (size_t(2) << log2len)) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->fft_mtx);
    DeviceGuard g(ctx->device);
    FftArgs a;
    rc = fft_args(ctx, direction, log2len, channels, d_in, d_out, stride, a);
    if (rc != TANX_OK) return rc;
    TANX_CHECK(launch_fft((int)channels, a, ctx->upd));
    TANX_CHECK(hipStreamSynchronize(ctx->upd));
    return TANX_OK;
}

// output ranges (floats) a transform defines in each channel buffer
static void fft_out_ranges(uint32_t dir, uint32_t log2len, size_t r[2][2])
{
    const size_t n = size_t(1) << log2len, m = n / 2;
    r[1][0] = r[1][1] = 0;
    switch (dir) {
    case TANX_FFT_FORWARD:
    case TANX_FFT_BACKWARD: r[0][0] = 0; r[0][1] = 2 * n; break;
    case TANX_FFT_R2C_FORWARD: r[0][0] = 0; r[0][1] = n + 2; break;
    case TANX_FFT_R2C_PLANAR_FORWARD:
        r[0][0] = 0; r[0][1] = m + 1;
        r[1][0] = m + 8; r[1][1] = 2 * m + 9;
        break;
    default: r[0][0] = 0; r[0][1] = n; break;
    }
}

extern "C" int32_t tanx_fft_transform(tanx_ctx *ctx, uint32_t direction, uint32_t log2len,
                                      uint32_t channels, const float *const *in, float *const *out)
{
    if (!ctx) return TANX_NOT_INITIALIZED;
    if (!in || !out) return TANX_INVALID_ARG;
    int32_t rc = fft_check(direction, log2len, channels);
    if (rc != TANX_OK) return rc;
    std::lock_guard<std::mutex> lk(ctx->fft_mtx);
    DeviceGuard g(ctx->device);
    const size_t stride = size_t(2) << log2len;
    const size_t need = stride * channels;
    if (ctx->fft_dev_floats < need) {
        if (ctx->fft_dev) (void)hipFree(ctx->fft_dev);
        ctx->fft_dev = nullptr;
        ctx->fft_dev_floats = 0;
        TANX_CHECK(hipMalloc(&ctx->fft_dev, need * sizeof(float)));
        ctx->fft_dev_floats = need;
    }
    if (ctx->fft_host_floats < need) {
        if (ctx->fft_host) (void)hipHostFree(ctx->fft_host);
        ctx->fft_host = nullptr;
        ctx->fft_host_floats = 0;
        TANX_CHECK(hipHostMalloc(&ctx->fft_host, need * sizeof(float)));
        ctx->fft_host_floats = need;
    }
    for (uint32_t c = 0; c < channels; c++) {
        if (!in[c] || !out[c]) return TANX_INVALID_ARG;
        std::memcpy(ctx->fft_host + c * stride, in[c], stride * sizeof(float));
    }
    hipStream_t s = ctx->upd;
    FftArgs a;
    rc = fft_args(ctx, direction, log2len, channels, ctx->fft_dev, ctx->fft_dev, stride, a);
    if (rc != TANX_OK) return rc;
    TANX_CHECK(hipMemcpyAsync(ctx->fft_dev, ctx->fft_host, need * sizeof(float), hipMemcpyHostToDevice, s));
    TANX_CHECK(launch_fft((int)channels, a, s));
    TANX_CHECK(hipMemcpyAsync(ctx->fft_host, ctx->fft_dev, need * sizeof(float), hipMemcpyDeviceToHost, s));
    TANX_CHECK(hipStreamSynchronize(s));
    size_t r[2][2];
    fft_out_ranges(direction, log2len, r);
    for (uint32_t c = 0; c < channels; c++)
        for (int k = 0; k < 2; k++)
            if (r[k][1] > r[k][0])
                std::memcpy(out[c] + r[k][0], ctx->fft_host + c * stride + r[k][0],
                            (r[k][1] - r[k][0]) * sizeof(float));
    return TANX_OK;
}

extern "C" const char *tanx_version(void) { return "tan_mi355x 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------
// output chain: TANMixer and TANConverter on the device
// ---------------------------------------------------------------------------
struct tanx_mixer {
    tanx_ctx *ctx = nullptr;
    std::mutex mtx;
    size_t n = 0;
    int C = 0;
    // device copies of pointer lists (ring) and their pinned staging
    const float **d_ptrs = nullptr;   // [kRunRing][C]
    const float **h_ptrs = nullptr;   // [kRunRing][C]
    hipEvent_t ev[kRunRing] = {};
    int ring = 0;
};

// host staging of at least `bytes` (device and pinned); the caller holds io_mtx
static int32_t io_staging(tanx_ctx *ctx, size_t bytes)
{
    if (!ctx->io_flag) {
        TANX_CHECK(hipMalloc(&ctx->io_flag, 16));
        TANX_CHECK(hipHostMalloc(&ctx->io_flag_host, 16));
    }
    if (ctx->io_cap >= bytes) return TANX_OK;
    TANX_CHECK(hipStreamSynchronize(ctx->proc));
    if (ctx->io_dev) (void)hipFree(ctx->io_dev);
    if (ctx->io_host) (void)hipHostFree(ctx->io_host);
    ctx->io_dev = nullptr;
    ctx->io_host = nullptr;
    ctx->io_cap = 0;
    size_t cap = 4096;
    while (cap < bytes) cap *= 2;
    TANX_CHECK(hipMalloc(&ctx->io_dev, cap));
    TANX_CHECK(hipHostMalloc(&ctx->io_host, cap));
    ctx->io_cap = cap;
    return TANX_OK;
}

/* TANMixer::Init (MixerImpl.cpp Init(buffer_size, num_channels)). */
extern "C" int32_t tanx_mixer_create(tanx_ctx *ctx, size_t buffer_size, int32_t channels, tanx_mixer **out)
{
    if (!ctx || !out) return TANX_INVALID_POINTER;
    *out = nullptr;
    if (buffer_size == 0 || channels <= 0) return TANX_INVALID_ARG;
    DeviceGuard g(ctx->device);
    if (!g.ok) return TANX_NO_DEVICE;
    tanx_mixer *m = new (std::nothrow) tanx_mixer();
    if (!m) return TANX_OUT_OF_MEMORY;
    m->ctx = ctx;
    ctx_retain(ctx);
    m->n = buffer_size;
    m->C = channels;
    hipError_t e = hipMalloc(&m->d_ptrs, sizeof(float *) * channels * kRunRing);
    if (e == hipSuccess) e = hipHostMalloc(&m->h_ptrs, sizeof(float *) * channels * kRunRing);
    for (auto &ev : m->ev)
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    for (auto &ev : m->ev)
        if (e == hipSuccess) e = hipEventRecord(ev, ctx->proc);
    if (e != hipSuccess) {
        if (m->d_ptrs) (void)hipFree(m->d_ptrs);
        if (m->h_ptrs) (void)hipHostFree(m->h_ptrs);
        for (auto &ev : m->ev) if (ev) (void)hipEventDestroy(ev);
        delete m;
        ctx_release(ctx);
        return map_hip(e);
    }
    *out = m;
    return TANX_OK;
}

extern "C" int32_t tanx_mixer_destroy(tanx_mixer *m)
{
    if (!m) return TANX_INVALID_POINTER;
    DeviceGuard g(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->proc);
    if (m->d_ptrs) (void)hipFree(m->d_ptrs);
    if (m->h_ptrs) (void)hipHostFree(m->h_ptrs);
    for (auto &ev : m->ev) if (ev) (void)hipEventDestroy(ev);
    tanx_ctx *ctx = m->ctx;
    delete m;
    ctx_release(ctx);
    return TANX_OK;
}

/* TANMixer::Mix(cl_mem pBufferInput[], cl_mem pBufferOutput)
 * (MixerImpl.cpp): d_in[c] are device buffers of buffer_size floats.
 * Asynchronous on the context's process stream. */
extern "C" int32_t tanx_mix_device(tanx_mixer *m, const float *const *d_in, float *d_out)
{
    if (!m) return TANX_NOT_INITIALIZED;
    if (!d_in || !d_out) return TANX_INVALID_ARG;
    for (int c = 0; c < m->C; c++)
        if (!d_in[c]) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> lk(m->mtx);
    DeviceGuard g(m->ctx->device);
    const int slot = m->ring;
    m->ring = (m->ring + 1) % kRunRing;
    TANX_CHECK(hipEventSynchronize(m->ev[slot]));
    const float **h = m->h_ptrs + (size_t)slot * m->C;
    const float **d = m->d_ptrs + (size_t)slot * m->C;
    for (int c = 0; c < m->C; c++) h[c] = d_in[c];
    hipStream_t s = m->ctx->proc;
    TANX_CHECK(hipMemcpyAsync(d, h, sizeof(float *) * m->C, hipMemcpyHostToDevice, s));
    TANX_CHECK(hipEventRecord(m->ev[slot], s));
    TANX_CHECK(launch_mix(d, nullptr, 0, m->C, (long long)m->n, d_out, s));
    return TANX_OK;
}

/* TANMixer::Mix(cl_mem pBufferInput, cl_mem pBufferOutput, amf_size inputStride)
 * (MixerImpl.cpp): channel c at d_in + c * in_stride.  Asynchronous. */
extern "C" int32_t tanx_mix_device_strided(tanx_mixer *m, const float *d_in, size_t in_stride, float *d_out)
{
    if (!m) return TANX_NOT_INITIALIZED;
    if (!d_in || !d_out || (m->C > 1 && in_stride < m->n)) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> lk(m->mtx);
    DeviceGuard g(m->ctx->device);
    TANX_CHECK(launch_mix(nullptr, d_in, (long long)in_stride, m->C, (long long)m->n, d_out, m->ctx->proc));
    return TANX_OK;
}

/* TANMixer::Mix(float* ppBufferInput[], float* ppBufferOutput) (MixerImpl.cpp):
 * host buffers; returns with the output written. */
extern "C" int32_t tanx_mix(tanx_mixer *m, const float *const *in, float *out)
{
    if (!m) return TANX_NOT_INITIALIZED;
    if (!in || !out) return TANX_INVALID_ARG;
    for (int c = 0; c < m->C; c++)
        if (!in[c]) return TANX_INVALID_ARG;
    tanx_ctx *ctx = m->ctx;
    std::lock_guard<std::mutex> lk(m->mtx);
    std::lock_guard<std::mutex> lk2(ctx->io_mtx);
    DeviceGuard g(ctx->device);
    const size_t n = m->n, bytes = (size_t)(m->C + 1) * n * sizeof(float);
    int32_t rc = io_staging(ctx, bytes);
    if (rc != TANX_OK) return rc;
    hipStream_t s = ctx->proc;
    TANX_CHECK(hipStreamSynchronize(s));  // staging free
    float *h = reinterpret_cast<float *>(ctx->io_host), *d = reinterpret_cast<float *>(ctx->io_dev);
    for (int c = 0; c < m->C; c++) std::memcpy(h + (size_t)c * n, in[c], n * sizeof(float));
    TANX_CHECK(hipMemcpyAsync(d, h, (size_t)m->C * n * sizeof(float), hipMemcpyHostToDevice, s));
    TANX_CHECK(launch_mix(nullptr, d, (long long)n, m->C, (long long)n, d + (size_t)m->C * n, s));
    TANX_CHECK(hipMemcpyAsync(h + (size_t)m->C * n, d + (size_t)m->C * n, n * sizeof(float), hipMemcpyDeviceToHost, s));
    TANX_CHECK(hipStreamSynchronize(s));
    std::memcpy(out, h + (size_t)m->C * n, n * sizeof(float));
    return TANX_OK;
}

static int sample_bytes(int32_t type) { return type == TANX_SAMPLE_SHORT ? 2 : 4; }

/* TANConverter::Convert(cl_mem, inputStep, inputOffset, inputType, cl_mem,
 * outputStep, outputOffset, outputType, n, gain, outputClipped)
 * (ConverterImpl.cpp:458-576): element i read at in[i * in_step + in_offset],
 * written at out[i * out_step + out_offset].  Asynchronous on the process
 * stream unless clipped is given: then it waits and reports whether
 * float -> short clamped any sample.  Returns AMF_OK when clipping happened,
 * as the reference does. */
extern "C" int32_t tanx_convert_device(tanx_ctx *ctx, const void *d_in, size_t in_step, size_t in_offset,
                                       int32_t in_type, void *d_out, size_t out_step, size_t out_offset,
                                       int32_t out_type, size_t n, float gain, int32_t *clipped)
{
    if (!ctx) return TANX_NOT_INITIALIZED;
    if (!d_in || !d_out || n == 0 || !(gain > 0.f)) return TANX_INVALID_ARG;
    if ((in_type != TANX_SAMPLE_FLOAT && in_type != TANX_SAMPLE_SHORT) ||
        (out_type != TANX_SAMPLE_FLOAT && out_type != TANX_SAMPLE_SHORT))
        return TANX_NOT_IMPLEMENTED;
    std::lock_guard<std::mutex> lk(ctx->io_mtx);
    DeviceGuard g(ctx->device);
    int32_t rc = io_staging(ctx, 0);
    if (rc != TANX_OK) return rc;
    hipStream_t s = ctx->proc;
    const bool flag = clipped && in_type == TANX_SAMPLE_FLOAT && out_type == TANX_SAMPLE_SHORT;
    if (flag) TANX_CHECK(launch_zero(ctx->io_flag, 16, s));
    TANX_CHECK(launch_convert(d_in, (long long)in_step, (long long)in_offset, in_type, d_out, (long long)out_step,
                              (long long)out_offset, out_type, (long long)n, gain, flag ? ctx->io_flag : nullptr, s));
    if (clipped) {
        *clipped = 0;
        if (flag) {
            TANX_CHECK(hipMemcpyAsync(ctx->io_flag_host, ctx->io_flag, sizeof(int), hipMemcpyDeviceToHost, s));
            TANX_CHECK(hipStreamSynchronize(s));
            *clipped = *ctx->io_flag_host ? 1 : 0;
        }
    }
    return TANX_OK;
}

/* TANConverter::Convert on host buffers (ConverterImpl.cpp:217-453): the
 * same conversion as the device path, run on the device through staging;
 * returns with the output written. */
extern "C" int32_t tanx_convert(tanx_ctx *ctx, const void *in, size_t in_step, int32_t in_type, void *out,
                                size_t out_step, int32_t out_type, size_t n, float gain, int32_t *clipped)
{
    if (!ctx) return TANX_NOT_INITIALIZED;
    if (!in || !out || n == 0 || !(gain > 0.f) || in_step == 0 || out_step == 0) return TANX_INVALID_ARG;
    if ((in_type != TANX_SAMPLE_FLOAT && in_type != TANX_SAMPLE_SHORT) ||
        (out_type != TANX_SAMPLE_FLOAT && out_type != TANX_SAMPLE_SHORT))
        return TANX_NOT_IMPLEMENTED;
    const size_t ib = sample_bytes(in_type), ob = sample_bytes(out_type);
    const size_t in_span = ((n - 1) * in_step + 1) * ib, out_span = ((n - 1) * out_step + 1) * ob;
    const size_t out_off = (in_span + 15) & ~size_t(15);
    std::lock_guard<std::mutex> lk(ctx->io_mtx);
    DeviceGuard g(ctx->device);
    int32_t rc = io_staging(ctx, out_off + out_span);
    if (rc != TANX_OK) return rc;
    hipStream_t s = ctx->proc;
    TANX_CHECK(hipStreamSynchronize(s));  // staging free
    char *h = ctx->io_host, *d = ctx->io_dev;
    std::memcpy(h, in, in_span);
    // strided outputs keep the samples between the written ones
    std::memcpy(h + out_off, out, out_span);
    TANX_CHECK(hipMemcpyAsync(d, h, out_off + out_span, hipMemcpyHostToDevice, s));
    const bool flag = in_type == TANX_SAMPLE_FLOAT && out_type == TANX_SAMPLE_SHORT;
    if (flag) TANX_CHECK(launch_zero(ctx->io_flag, 16, s));
    TANX_CHECK(launch_convert(d, (long long)in_step, 0, in_type, d + out_off, (long long)out_step, 0, out_type,
                              (long long)n, gain, flag ? ctx->io_flag : nullptr, s));
    TANX_CHECK(hipMemcpyAsync(h + out_off, d + out_off, out_span, hipMemcpyDeviceToHost, s));
    if (flag) TANX_CHECK(hipMemcpyAsync(ctx->io_flag_host, ctx->io_flag, sizeof(int), hipMemcpyDeviceToHost, s));
    TANX_CHECK(hipStreamSynchronize(s));
    std::memcpy(out, h + out_off, out_span);
    if (clipped) *clipped = flag && *ctx->io_flag_host ? 1 : 0;
    return TANX_OK;
}

// ---------------------------------------------------------------------------
// block DFT on the matrix cores
// ---------------------------------------------------------------------------
static uint16_t bf16_rne(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
static float bf16_to_f(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// tables of the n-point real DFT: column 2k = cos(2 pi k t / n), column 2k+1
// = -sin(2 pi k t / n), values in FP64 split into bf16 high + low parts
static int32_t dft_tables(tanx_ctx *ctx, int n)
{
    if (ctx->dft_n == n) return TANX_OK;
    TANX_CHECK(hipStreamSynchronize(ctx->proc));
    if (ctx->dft_hi) (void)hipFree(ctx->dft_hi);
    if (ctx->dft_lo) (void)hipFree(ctx->dft_lo);
    ctx->dft_hi = ctx->dft_lo = nullptr;
    ctx->dft_n = 0;
    const size_t cols = (size_t)n + 2, count = cols * n;
    std::vector<uint16_t> hi(count), lo(count);
    for (size_t col = 0; col < cols; col++) {
        const size_t k = col / 2;
        for (int t = 0; t < n; t++) {
            const double ang = 2.0 * M_PI * (double)((k * (size_t)t) % (size_t)n) / (double)n;
            const double v = (col & 1) ? -std::sin(ang) : std::cos(ang);
            const uint16_t h = bf16_rne((float)v);
            hi[col * n + t] = h;
            lo[col * n + t] = bf16_rne((float)(v - (double)bf16_to_f(h)));
        }
    }
    TANX_CHECK(hipMalloc(&ctx->dft_hi, count * 2));
    TANX_CHECK(hipMalloc(&ctx->dft_lo, count * 2));
    // ordered on the process stream (the kernels' stream) and complete before
    // the host vectors go away
    TANX_CHECK(hipMemcpyAsync(ctx->dft_hi, hi.data(), count * 2, hipMemcpyHostToDevice, ctx->proc));
    TANX_CHECK(hipMemcpyAsync(ctx->dft_lo, lo.data(), count * 2, hipMemcpyHostToDevice, ctx->proc));
    TANX_CHECK(hipStreamSynchronize(ctx->proc));
    ctx->dft_n = n;
    return TANX_OK;
}

/* Real-to-complex DFT of 2^log2len samples per channel on the matrix cores
 * (the block DFT of the precision track; no counterpart in the reference,
 * whose R2C transforms are FFTs: FFTImpl.cpp:996-1122).  Output as
 * TANFFT's R2C direction: bins 0..n/2 interleaved (re, im). */
extern "C" int32_t tanx_dft_r2c_mfma_device(tanx_ctx *ctx, uint32_t log2len, uint32_t channels, const float *d_in,
                                            size_t in_stride, float *d_out, size_t out_stride, uint32_t split)
{
    if (!ctx) return TANX_NOT_INITIALIZED;
    if (!d_in || !d_out || channels == 0 || log2len < 5 || log2len > 11 || (split != 1 && split != 3))
        return TANX_INVALID_ARG;
    const int n = 1 << log2len;
    if (in_stride < (size_t)n || out_stride < (size_t)n + 2) return TANX_INVALID_ARG;
    // the kernel reads 16-byte vectors of samples
    if ((in_stride & 3) || (reinterpret_cast<uintptr_t>(d_in) & 15)) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->dft_mtx);
    DeviceGuard g(ctx->device);
    int32_t rc = dft_tables(ctx, n);
    if (rc != TANX_OK) return rc;
    TANX_CHECK(launch_dft_mfma(d_in, (long long)in_stride, d_out, (long long)out_stride, (int)channels, n,
                               ctx->dft_hi, ctx->dft_lo, (int)split, ctx->proc));
    return TANX_OK;
}

// ---------------------------------------------------------------------------
// IR sources: room responses by image sources
// ---------------------------------------------------------------------------

// Derives the ears of `count` responses, grows the context's buffers and
// launches; the caller holds room_mtx and the device guard.
static int32_t room_launch(tanx_ctx *ctx, uint32_t count, const tanx_room *rooms, const tanx_source *sources,
                           const tanx_listener *listeners, int32_t fs, int32_t len, uint32_t flags, int32_t max_bounces,
                           float *d_out, size_t out_stride)
{
    const size_t nears = 2 * (size_t)count;
    std::vector<RoomEar> ears(nears);
    long long max_early = 0;
    int max_windows = 0;
    for (uint32_t k = 0; k < count; k++) {
        long long images = 0;
        const int32_t rc = room_ears(&rooms[k], &sources[k], &listeners[k], fs, len, flags, max_bounces,
                                     &ears[2 * k], &images);
        if (rc != TANX_OK) return rc;
        for (int e = 0; e < 2; e++) {
            const RoomEar &E = ears[2 * k + e];
            long long box = 1;
            for (int a = 0; a < 3; a++) box *= E.hi[a] >= E.lo[a] ? (long long)(E.hi[a] - E.lo[a] + 1) : 0;
            max_early = std::max(max_early, box);
            max_windows = std::max(max_windows, (len - E.wstart + kRoomWindow - 1) / kRoomWindow);
        }
    }
    const size_t ints = nears * (size_t)len;
    if (ints > ctx->room_acc_ints || nears > ctx->room_ears_cap) {
        // buffers in use by earlier calls on the stream: let them finish
        TANX_CHECK(hipStreamSynchronize(ctx->proc));
    }
    if (ints > ctx->room_acc_ints) {
        if (ctx->room_acc) (void)hipFree(ctx->room_acc);
        ctx->room_acc = nullptr;
        ctx->room_acc_ints = 0;
        TANX_CHECK(hipMalloc(&ctx->room_acc, ints * sizeof(unsigned)));
        ctx->room_acc_ints = ints;
        TANX_CHECK(launch_zero(ctx->room_acc, ints * sizeof(unsigned), ctx->proc));
    }
    if (nears > ctx->room_ears_cap) {
        if (ctx->room_ears_dev) (void)hipFree(ctx->room_ears_dev);
        if (ctx->room_ears_host) (void)hipHostFree(ctx->room_ears_host);
        ctx->room_ears_dev = ctx->room_ears_host = nullptr;
        ctx->room_ears_cap = 0;
        ctx->room_ev_live = false;
        TANX_CHECK(hipMalloc(&ctx->room_ears_dev, nears * sizeof(RoomEar)));
        TANX_CHECK(hipHostMalloc(&ctx->room_ears_host, nears * sizeof(RoomEar), hipHostMallocDefault));
        ctx->room_ears_cap = nears;
    }
    if (!ctx->room_ev) TANX_CHECK(hipEventCreateWithFlags(&ctx->room_ev, hipEventDisableTiming));
    // the previous call's staging copy must be done before it is rewritten
    if (ctx->room_ev_live) TANX_CHECK(hipEventSynchronize(ctx->room_ev));
    std::memcpy(ctx->room_ears_host, ears.data(), nears * sizeof(RoomEar));
    TANX_CHECK(hipMemcpyAsync(ctx->room_ears_dev, ctx->room_ears_host, nears * sizeof(RoomEar),
                              hipMemcpyHostToDevice, ctx->proc));
    TANX_CHECK(hipEventRecord(ctx->room_ev, ctx->proc));
    ctx->room_ev_live = true;
    TANX_CHECK(launch_room(ctx->room_ears_dev, (int)nears, max_early, max_windows, len, ctx->room_acc, d_out,
                           (long long)out_stride, ctx->proc));
    return TANX_OK;
}

extern "C" int32_t tanx_room_response_device(tanx_ctx *ctx, uint32_t count, const tanx_room *rooms,
                                             const tanx_source *sources, const tanx_listener *listeners,
                                             int32_t sample_rate, int32_t response_length, uint32_t flags,
                                             int32_t max_bounces, float *d_out, size_t out_stride)
{
    if (!ctx) return TANX_NOT_INITIALIZED;
    if (!rooms || !sources || !listeners || !d_out) return TANX_INVALID_POINTER;
    if (count == 0 || count > 32767 || response_length <= 0 || out_stride < (size_t)response_length)
        return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->room_mtx);
    DeviceGuard g(ctx->device);
    return room_launch(ctx, count, rooms, sources, listeners, sample_rate, response_length, flags, max_bounces, d_out,
                       out_stride);
}

extern "C" int32_t tanx_room_response(tanx_ctx *ctx, const tanx_room *room, const tanx_source *source,
                                      const tanx_listener *listener, int32_t sample_rate, int32_t response_length,
                                      uint32_t flags, int32_t max_bounces, float *left, float *right)
{
    if (!ctx) return TANX_NOT_INITIALIZED;
    if (!room || !source || !listener || !left || !right) return TANX_INVALID_POINTER;
    if (response_length <= 0) return TANX_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->room_mtx);
    DeviceGuard g(ctx->device);
    const size_t need = 2 * (size_t)response_length;
    if (need > ctx->room_out_cap) {
        TANX_CHECK(hipStreamSynchronize(ctx->proc));
        if (ctx->room_out_dev) (void)hipFree(ctx->room_out_dev);
        if (ctx->room_out_host) (void)hipHostFree(ctx->room_out_host);
        ctx->room_out_dev = ctx->room_out_host = nullptr;
        ctx->room_out_cap = 0;
        TANX_CHECK(hipMalloc(&ctx->room_out_dev, need * sizeof(float)));
        TANX_CHECK(hipHostMalloc(&ctx->room_out_host, need * sizeof(float), hipHostMallocDefault));
        ctx->room_out_cap = need;
    }
    int32_t rc = room_launch(ctx, 1, room, source, listener, sample_rate, response_length, flags, max_bounces,
                             ctx->room_out_dev, (size_t)response_length);
    if (rc != TANX_OK) return rc;
    TANX_CHECK(hipMemcpyAsync(ctx->room_out_host, ctx->room_out_dev, need * sizeof(float), hipMemcpyDeviceToHost,
                              ctx->proc));
    TANX_CHECK(hipStreamSynchronize(ctx->proc));
    std::memcpy(left, ctx->room_out_host, (size_t)response_length * sizeof(float));
    std::memcpy(right, ctx->room_out_host + response_length, (size_t)response_length * sizeof(float));
    return TANX_OK;
}
