// Internal declarations shared by the HIP kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

// public parameter structs of the room-response entry points (tan_mi355x.h)
struct tanx_room;
struct tanx_source;
struct tanx_listener;

namespace tanx {

// Spectra are stored "packed": a 2B-point real FFT of a frame of 2B reals
// has B+1 bins whose DC and Nyquist bins are real; element 0 holds
// (DC.re, Nyquist.re) and elements 1..B-1 hold bins 1..B-1.  Every spectrum
// row is therefore exactly B complex floats (B*8 bytes).

// One launch of k_tick runs up to two roles:
//   head (blocks [0, nhead)): channel c = block (heads-only group launches:
//     c = wave_heads(B) * block + wave); R2C of [x_{t-1}, x_t], FDL insert,
//     A_f = sum of tail partials + X_{t-1} H_{f,1} + X_t H_{f,0}, C2R,
//     crossfade, output;
//   tail (blocks [nhead, nhead + C * nchunk)): partial sums over partitions
//     [2 + j*pc, 2 + (j+1)*pc) of X_{t+1-i} H_i for the NEXT block;
//   frame multiply-accumulate (blocks after those, non-uniform objects only,
//     C * nchB of them): bins [j*B, (j+1)*B) of the frame segment's partial
//     sums over frame partitions [2 + k*pcB, 2 + (k+1)*pcB).
// The roles touch disjoint data, so they may share a launch.
//
// Non-uniform partitioned objects (frame N = M * B samples): taps [0, N)
// run as the uniform engine above with P = M partitions of B; taps [N, len)
// form PB = len / N - 1 partitions of N, convolved frame by frame with
// 2N-point transforms.  Frame g (input samples [gN, (g+1)N)) is complete at
// block gM + M - 1; its output covers blocks (g+1)M .. (g+1)M + M - 1.  The
// sum over frame partitions 2..PB-1 only needs frames <= g - 2, so it is
// spread over blocks gM - 1 .. gM + M - 2, one slice of B bins per block;
// the frame kernel (FrameArgs) at block gM + M - 1 adds partitions 0 and 1
// and the slices, and inverse-transforms the frame's output.
// IR spectrum storage: FP32 complex, or FP16 / BF16 complex scaled per
// partition row (value = stored * S[c * P + i], S = the row's max |re|,|im|),
// always accumulated in FP32.
enum IrType { IR_F32 = 0, IR_F16 = 1, IR_BF16 = 2 };
inline int ir_elem_bytes(int t) { return t == IR_F32 ? 8 : 4; }

constexpr int kPassMax = 16;     // blocks one tail pass computes partial sums for, at most
constexpr int kPassDefault = 4;  // the library's choice for IRs of 16 or more partitions
constexpr int kHeadBatch = 4;    // heads one launch runs together, at most
// partitions a head adds itself at most: a pass of the launch with heads
// [w, e) serves blocks up to e + kPassMax - 1 from rows <= w - 1, e - w <= kHeadBatch
constexpr int kNearMax = kHeadBatch + kPassMax - 1;
// Heads per workgroup of a heads-only group launch (objects without a pass,
// tail or frame slice in a tick): one per wave for blocks of at most 256
// samples (an FFT stage of the head is at most 64 butterflies, one wave's
// worth), one per workgroup above.  Other tick launches run a head per
// workgroup.
constexpr int wave_heads(int B) { return B <= 256 ? 4 : 1; }

struct TickArgs {
    int B;                 // block size (group launches dispatch on it)
    int C;
    int P;
    int irtype;            // IrType of Ha/Hb/tHa/tHb
    const uint8_t *run;    // [C] 1 = channel running
    const float2 *twM;     // W_B^t, t < B
    const float2 *twN;     // W_2B^t, t < 2B
    float2 *fdl;           // [C][P][B] ring of input spectra
    int cached;            // nonzero: the streamed rows (FDL, IR) fit in the Infinity Cache and
                           // k_window's pass reads them with ordinary loads; 0: nontemporal loads

    // head role
    int nhead;             // C or 0
    const float *xin;      // channel c block at xin + c * in_stride
    long long in_stride;
    float *yout;           // channel c output at yout + c * out_stride
    long long out_stride;
    float *xprev;          // [C][B] previous input block (overlap-save frame)
    int slot;              // ring slot receiving X_t
    const void *Ha;        // filter slot bases [C][P][B] of IrType
    const void *Hb;        // crossfade: new filter, else nullptr
    const float *Sa;       // row scales [C][P] (reduced precision)
    const float *Sb;
    const float2 *Ta;      // partial sums read by the head [C][tchunk][B]
    const float2 *Tb;
    int tchunk;
    int xfade;             // 0: y = y_A; 1: y = y_B * j/F + y_A * (F-j)/F

    // tail role
    int nchunk;            // chunks per channel (0 = no tail blocks)
    int pc;                // partitions per chunk
    int tbase;             // ring slot of block t+1
    const void *tHa;
    const void *tHb;       // second filter (next block crossfades) or nullptr
    const float *tSa;
    const float *tSb;
    float2 *tTa;           // partial sums written [C][nchunk][B]
    float2 *tTb;
    // multi-block pass (kn > 1) from FDL rows before the launch's first head
    // block o: output k (0-based) holds the partial sums of block o + kgap + k
    // over partitions i >= kgap + k + 1 (FDL row (tbase + k - i) & (P - 1),
    // tbase the ring slot of block o + kgap); filter tHa's sums go to
    // kTA[k], filter tHb's to kTB[k] (nullptr: not computed).  tTa/tTb unused.
    int kp;                // the object's pass length (the launch compiles passes of at most 4 or 8 outputs)
    int kn;
    int kgap;
    float2 *kTA[kPassMax];
    float2 *kTB[kPassMax];
    // head: partitions 0..dA (filter A) and 0..dB (filter B), at most
    // kNearMax, are added by the head itself (1 without multi-block passes);
    // the rest come from Ta/Tb
    int dA, dB;

    // non-uniform objects (nu = 0 for uniform ones)
    int nu;
    int M;                 // blocks per frame
    int N;                 // frame length M * B
    float *xring;          // [C][2M][B] last 2M input blocks
    int xslot;             // ring slot receiving x_t (t mod 2M)
    const float *zA;       // frame output added to filter A's y: [C][N] at + zoff, or nullptr
    const float *zB;       // the same for filter B (crossfade blocks)
    int zoff;              // (t mod M) * B

    // frame multiply-accumulate role
    int nchB;              // chunks per channel (0 = none this block)
    int pcB;               // frame partitions per chunk
    int PB;                // frame partitions
    int RB;                // frame ring rows (power of two > PB)
    const float2 *fb;      // [C][RB][N] frame spectra ring
    int fbase;             // ring row of the frame being summed for
    int bslice;            // bin slice (bins bslice*B .. +B)
    const void *bHa;       // frame-partition spectra [C][PB][N] of IrType
    const void *bHb;       // second filter or nullptr
    const float *bSa;      // row scales [C][PB]
    const float *bSb;
    float2 *bTa;           // partial sums written [C][nchB][N] (slice bins)
    float2 *bTb;
};

// The tail role of a launch is a multi-block pass (tailk_role: outputs in
// kTA/kTB, any kn >= 1, partitions from kgap + k + 1) rather than a single
// tail (tail_role: outputs in tTa/tTb, partitions from 2).  A pass may be one
// block long: one that ends on its object's pass-phase grid.
__host__ __device__ inline bool is_pass(const TickArgs &a)
{
    return a.kn > 1 || a.kTA[0] != nullptr || a.kTB[0] != nullptr;
}

// Pass length a launch compiles for an object of pass length kp (registers
// grow with it): 4, 8, 12 or 16.
__host__ __device__ inline int pass_km(int kp) { return kp > 12 ? 16 : kp > 8 ? 12 : kp > 4 ? 8 : 4; }

// Time-domain convolution (TAN_CONVOLUTION_METHOD_TIME_DOMAIN): one
// workgroup per channel computes n <= B outputs
// y[j] = sum_{k=first}^{last-1} hist[(pos + j - k) mod len] h[k] over the
// channel's circular history (the new samples written at pos..pos+n-1), with
// a one-block crossfade to the second filter when xfade is set, and then
// writes the new samples into the history.
constexpr int kMaxTdLen = 4096;  // window and two filters within 64 KB of LDS at B = 1024
struct TdArgs {
    int B, C, len, n;
    const uint8_t *run;
    const float *xin;
    long long in_stride;
    float *yout;
    long long out_stride;
    float *hist;           // [C][len]
    int pos;               // history position of the first new sample
    const float *Ha;       // filters: channel c at Ha + c * h_stride
    const float *Hb;       // crossfade: new filter, else nullptr
    long long h_stride;
    const int *nz;         // [C][2] first / last+1 non-zero tap, or nullptr (0, len)
    int xfade;
};

// Frame kernel of a non-uniform object at the last block of frame g: one
// workgroup per channel.  X_g = R2C_2N([x_{g-1}, x_g]) from the input ring
// (the head of the same block has written x_t into it), into the frame ring;
// for each filter f: A = X_g H_f0 + X_{g-1} H_f1 + sum of the frame partials,
// z_f = C2R_2N(A)[N, 2N) / 2N, the frame's output.
struct FrameArgs {
    int B, M, N, C, PB, RB;
    const uint8_t *run;
    const float *xring;    // [C][2M][B]
    int xlast;             // ring slot of the frame's last block
    float2 *fb;            // [C][RB][N]
    int frow;              // ring row of X_g
    const void *Ha;        // frame-partition spectra [C][PB][N] of IrType
    const void *Hb;        // second filter (the next block crossfades) or nullptr
    const float *Sa;
    const float *Sb;
    const float2 *Ta;      // frame partials [C][nch][N] (nch may be 0)
    const float2 *Tb;
    int nch;
    float *za;             // [C][N] outputs
    float *zb;
    const float2 *twN;     // W_N^t, t < N
    const float2 *tw2N;    // W_2N^t, t < 2N
};

struct IrArgs {
    const float *ir;       // channel ch IR at ir + ch * stride, n valid samples
    long long stride;
    long long n;
    const int *chan;       // list of channels to transform (blockIdx.y indexes it); null: channel blockIdx.y
    void *H;               // destination slot base [C][P][B] of IrType
    float *S;              // destination row scales [C][P]
    int irtype;
    int P;
    const float2 *twM;
    const float2 *twN;
};

struct FftArgs {
    const float *in;
    float *out;
    long long stride;      // floats between channels
    int mode;              // TAN_FFT_TRANSFORM_DIRECTION
    int log2len;
    const float2 *twM;     // W_M table for the complex FFT size M
    const float2 *twN;     // W_2M table (real transforms)
    // complex sizes M > kFftLdsMax: the four-step transform (M = M1 * M2,
    // columns of M1 points, then rows of M2 points) through scratch of
    // 2 * M complex per channel, with the W_M1 and W_M2 tables
    float2 *scratch;
    const float2 *tw1, *tw2;
};
constexpr int kFftLdsMax = 2048;   // complex points one workgroup transforms in LDS
constexpr int kFftLog2Max = 19;    // largest TANFFT log2len (complex and real)

// Launchers (return hipError_t).  B is the block size (power of two).
// ev0/ev1 (optional): events stamped with the kernel's own start and end
hipError_t launch_tick(int B, int tail_nf, const TickArgs &a, hipStream_t s, hipEvent_t ev0 = nullptr,
                       hipEvent_t ev1 = nullptr);
// One launch for several objects of block size B: block b belongs to the
// object o with starts[o] <= b < starts[o + 1] and runs its roles from
// d_descs[o].  Every object must use irtype, block size B (64, 128, 256 or
// 512), role set `roles` (group_roles) and, for passes, pass_km(kp) == km.
// IR transforms of several objects of partition size B and one IR type in
// one launch: entry j (d_descs[j], its channel list of nchan_j entries) owns
// blocks [d_starts[j], d_starts[j] + ir_transform_blocks(B, irtype, P_j,
// nchan_j)).
int ir_transform_blocks(int B, int irtype, int P, int nchan);
hipError_t launch_ir_transform_group(int B, int irtype, const IrArgs *d_descs, const int *d_starts, int n,
                                     unsigned blocks, hipStream_t s);
hipError_t launch_group_tick(int irtype, int B, int roles, int km, const TickArgs *d_descs, const int *d_starts,
                             int nobj, unsigned blocks, hipStream_t s, hipEvent_t ev0 = nullptr,
                             hipEvent_t ev1 = nullptr);
// Work-queue form of launch_group_tick: `total` workgroup items taken in
// order from *next (zero before the launch) by the device's resident
// workgroups of the kernel.
hipError_t launch_group_queue(int irtype, int B, int roles, int km, const TickArgs *d_descs, const int *d_starts,
                              int nobj, int total, int *next, hipStream_t s, hipEvent_t ev0 = nullptr,
                              hipEvent_t ev1 = nullptr);
// Role set of a group descriptor (kRoles*): heads with a single-filter pass
// (lean), with a two-filter pass (crossfade), with a single tail or a frame
// slice (general), or heads alone (heads: wave_heads(B) per workgroup).
constexpr int kRolesGeneral = 0, kRolesLean = 1, kRolesXfade = 2, kRolesHeads = 3;
__host__ __device__ inline int group_roles(const TickArgs &a)
{
    if (a.nchB > 0 || (a.nchunk > 0 && !is_pass(a))) return kRolesGeneral;
    if (a.nchunk == 0) return kRolesHeads;
    return a.tHb != nullptr ? kRolesXfade : kRolesLean;
}
// Workgroups of a group descriptor's roles in a launch of role set `roles`
inline int group_blocks(const TickArgs &a, int roles)
{
    const int heads = roles == kRolesHeads ? (a.nhead + wave_heads(a.B) - 1) / wave_heads(a.B) : a.nhead;
    return heads + a.C * (a.nchunk + a.nchB);
}
constexpr int kGroupMaxB = 512;
// Consecutive head-only blocks of a uniform object with multi-block passes,
// launched together (one workgroup per channel): block j of the batch
// follows block j-1; it adds its partial sums T[j] (tchunk[j] rows per
// channel) and partitions 0..d[j] (d[j] <= kNearMax) of filter H[j] itself,
// writes its spectrum to FDL slot[j], and the last block's input becomes
// xprev.  No crossfade or frame role; B <= kHeadBatchMaxB.  launch_heads
// runs the batch alone; with a pass (TickArgs a with kn > 1 and its gap
// after the batch), launch_window runs the batch and the pass in one
// launch: the pass reads FDL rows before the batch only.
constexpr int kHeadBatchMaxB = 256;  // larger blocks: one head per launch (registers)
struct HeadBatchArgs {
    int C, P;
    int kp;                // the object's pass length: d[j] <= kHeadBatch + kp - 1
    const uint8_t *run;
    const float2 *twM, *twN;
    float2 *fdl;
    float *xprev;
    long long in_stride, out_stride;
    int n;
    const float *xin[kHeadBatch];
    float *yout[kHeadBatch];
    const void *H[kHeadBatch];
    const float *S[kHeadBatch];
    const float2 *T[kHeadBatch];
    int slot[kHeadBatch], tchunk[kHeadBatch], d[kHeadBatch];
};
hipError_t launch_heads(int B, int irtype, const HeadBatchArgs &h, hipStream_t s, hipEvent_t ev0 = nullptr,
                        hipEvent_t ev1 = nullptr);
hipError_t launch_window(int B, int tail_nf, const HeadBatchArgs &h, const TickArgs &a, hipStream_t s,
                         hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// k_tick reading its arguments from d_descs[k] (graph replay; kn must be 1)
hipError_t launch_tick_desc(int B, int irtype, const TickArgs *d_descs, int k, unsigned blocks, hipStream_t s);
hipError_t launch_ir_transform(int B, int nchan, const IrArgs &a, hipStream_t s);
// frame kernels: N in {128, ..., 4096}
constexpr int kMaxFrame = 4096;
hipError_t launch_frame(int irtype, const FrameArgs &a, hipStream_t s, hipEvent_t ev1 = nullptr);
hipError_t launch_td(const TdArgs &a, hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// strided copy of IR rows into a [C][len] slot, zero past n (list of channels)
hipError_t launch_td_load(const float *ir, long long stride, long long n, const int *chan, int nchan, float *H,
                          int len, hipStream_t s);
hipError_t launch_group_frame(int irtype, const FrameArgs *d_descs, const int *d_starts, int nobj, unsigned blocks,
                              hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_fft(int channels, const FftArgs &a, hipStream_t s);
// batched real DFT of n-sample frames on the matrix cores (n a multiple of
// 32): y[c][2k], y[c][2k+1] = Re, Im of bin k <= n/2; fh/fl: [n+2][n] bf16
// tables (column-major in t); split 1 or 3 products
hipError_t launch_dft_mfma(const float *x, long long in_stride, float *y, long long out_stride, int C, int n,
                           const void *fh, const void *fl, int split, hipStream_t s);
// output chain: channel sum, and sample conversion (types 0 = float, 1 = short)
hipError_t launch_mix(const float *const *d_ins, const float *d_in, long long stride, int channels, long long n,
                      float *d_out, hipStream_t s);
hipError_t launch_convert(const void *in, long long in_step, long long in_off, int in_type, void *out,
                          long long out_step, long long out_off, int out_type, long long n, float gain, int *d_flag,
                          hipStream_t s);
hipError_t launch_zero(void *p, size_t bytes, hipStream_t s);

// Room responses by image sources (GenerateRoomResponse.cl).  One RoomEar
// per response row: the kernel arguments the reference's host code derives
// per ear (TrueAudioVR.cpp:659-778).  Images ix in [-n[0], n[0]] (x: width),
// iy in [-n[1], n[1]] (y: height), iz in [-n[2], n[2]] (z: length).
constexpr int kHeadFilterMax = 64;      // SIMPLEHRTF_FFTLEN
constexpr float kRoomFixedOne = 67108864.0f;  // 2^26: one in the fixed-point accumulator
struct RoomEar {
    float src[3];      // source position, clamped into the room
    float head[3];     // ear position
    float ev[3];       // ear direction
    float ev_norm;     // |ev|
    float dim[3];      // width, height, length
    float damp[6];     // right, left, top, bottom, front, back
    float max_gain, d_min;
    int fs, len, hrtf_len, filt_len;
    int n[3];
    int suppress_direct;
    float lp[kHeadFilterMax], hp[kHeadFilterMax];
    // the box of image indices that can arrive before hrtf_len (the early,
    // head-filtered images): lo[k] .. hi[k] per axis
    int lo[3], hi[3];
    int wstart;        // first output sample summed by windows: min(hrtf_len, len)
};
constexpr int kRoomWindow = 2048;  // output samples per late-image window (8 KB of LDS)
// RoomEar of the left (out[0]) and right (out[1]) ear; AMF_RESULT value.
int32_t room_ears(const tanx_room *room, const tanx_source *source, const tanx_listener *ears, int fs, int len,
                  uint32_t flags, int max_bounces, RoomEar out[2], long long *images);
// Sums the images of ears[e] (wrapping 32-bit integer sums of the
// fixed-point contributions) into row e of out (out + e * out_stride, len
// floats).  Early images go into acc[e * len ...] with global atomics, late
// ones through per-window LDS sums added into acc; acc is left cleared.
// max_early: the largest early-box image count over the ears; max_windows:
// the largest number of late windows.
hipError_t launch_room(const RoomEar *d_ears, int nears, long long max_early, int max_windows, int len,
                       unsigned *acc, float *out, long long out_stride, hipStream_t s);

// Allocation-free group launches (tanx_group_process_device_noalloc): the
// descriptors of at most kNoAllocGroupMax objects per call are staged in a
// file-scope device array of kNoAllocScratchBytes (tan_noalloc.hip), laid out
// as [TickArgs x max][int x max][FrameArgs x max][int x max], each piece
// 16-byte aligned.
constexpr int kNoAllocGroupMax = 1024;
constexpr int kNoAllocMaxDevices = 64;
constexpr size_t align16(size_t b) { return (b + 15) & ~size_t(15); }
constexpr size_t kNoAllocDescBytes = align16(sizeof(TickArgs) * kNoAllocGroupMax);
constexpr size_t kNoAllocStartBytes = align16(sizeof(int) * kNoAllocGroupMax);
constexpr size_t kNoAllocFrameBytes = align16(sizeof(FrameArgs) * kNoAllocGroupMax);
constexpr size_t kNoAllocScratchBytes = kNoAllocDescBytes + 2 * kNoAllocStartBytes + kNoAllocFrameBytes;
// Resolves the scratch array's address on `device`, which must be current;
// once per device (later calls return at once).
hipError_t noalloc_scratch_init(int device);
// The address noalloc_scratch_init resolved for `device`, or nullptr.
void *noalloc_scratch(int device);

// Twiddle tables: tw + (2^l - 1) holds W_{2^l}^t = exp(-2 pi i t / 2^l), t < 2^l,
// for l = 0..kMaxTwLog2.
constexpr int kMaxTwLog2 = 13;
inline const float2 *tw_table(const float2 *tw, int log2) { return tw + ((1 << log2) - 1); }

}  // namespace tanx
