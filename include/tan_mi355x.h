/*
 * tan_mi355x.h -- C ABI of the MI355X (gfx950) TAN convolution library.
 *
 * The library implements TAN's time-varying multichannel convolution path
 * (uniformly partitioned FFT convolution with a frequency-domain delay line
 * and IR crossfade) and TANFFT on one GPU.  Every entry point takes plain
 * pointers and sizes and returns an int32_t equal to an AMF_RESULT value
 * (amf/public/include/core/Result.h:43-122).  Objects are thread-safe per
 * object, as the reference's are (locks at ConvolutionImpl.h:210-216).
 *
 * Reference paths below are relative to the TAN checkout.
 */
#ifndef TAN_MI355X_H
#define TAN_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tanx_ctx tanx_ctx;
typedef struct tanx_conv tanx_conv;

/* AMF_RESULT values used here (Result.h:43-122). */
enum {
    TANX_OK = 0,
    TANX_FAIL = 1,
    TANX_UNEXPECTED = 2,
    TANX_INVALID_ARG = 4,
    TANX_OUT_OF_RANGE = 5,
    TANX_OUT_OF_MEMORY = 6,
    TANX_INVALID_POINTER = 7,
    TANX_NOT_IMPLEMENTED = 9,
    TANX_NOT_SUPPORTED = 10,
    TANX_NOT_FOUND = 11,
    TANX_ALREADY_INITIALIZED = 12,
    TANX_NOT_INITIALIZED = 13,
    TANX_INVALID_FORMAT = 14,
    TANX_WRONG_STATE = 15,
    TANX_FILE_NOT_OPEN = 16,
    TANX_NO_DEVICE = 17,
    TANX_INPUT_FULL = 25,
    TANX_TAN_UNSUPPORTED_VERSION = 43
};

/* TAN_CONVOLUTION_METHOD (tan/tanlibrary/include/TrueAudioNext.h:81-98). */
enum {
    TANX_METHOD_FFT_OVERLAP_ADD = 0,
    TANX_METHOD_FFT_PARTITIONED_UNIFORM = 1,
    TANX_METHOD_FFT_PARTITIONED_NONUNIFORM = 2,
    TANX_METHOD_FFT_UNIFORM_PARTITIONED = 3,
    TANX_METHOD_FHT_UNIFORM_PARTITIONED = 4,
    TANX_METHOD_FFT_UNIFORM_HEAD_TAIL = 5,
    TANX_METHOD_FHT_UNIFORM_HEAD_TAIL = 6,
    TANX_METHOD_TIME_DOMAIN = 7,
    TANX_METHOD_FHT_NONUNIFORM_PARTITIONED = 8,
    TANX_METHOD_FFT_NONUNIFORM_PARTITIONED = 9,
    TANX_METHOD_USE_PROCESS_FINALIZE = 0x8000,
    TANX_METHOD_USE_PROCESS_TAILTHREAD = 0xC000
};

/* TAN_CONVOLUTION_CHANNEL_FLAG / TAN_CONVOLUTION_OPERATION_FLAG
 * (TrueAudioNext.h:105-120). */
enum {
    TANX_CHANNEL_FLAG_PROCESS = 0,
    TANX_CHANNEL_FLAG_STOP_INPUT = 0x01,
    TANX_CHANNEL_FLAG_FLUSH_STREAM = 0x02,
    TANX_OPERATION_FLAG_NONE = 0,
    TANX_OPERATION_FLAG_BLOCK_UNTIL_READY = 0x01
};

/* TAN_FFT_TRANSFORM_DIRECTION (TrueAudioNext.h:493-501). */
enum {
    TANX_FFT_FORWARD = 0,
    TANX_FFT_BACKWARD = 1,
    TANX_FFT_R2C_FORWARD = 2,
    TANX_FFT_C2R_BACKWARD = 3,
    TANX_FFT_R2C_PLANAR_FORWARD = 4,
    TANX_FFT_C2R_PLANAR_BACKWARD = 5
};

/* ---- context ------------------------------------------------------------ */

/* Replaces TANCreateContext + TANContext::InitOpenCL
 * (core/TANContextImpl.cpp:87-132, 264-328): binds HIP device `device` and
 * creates the two streams that stand in for the general (IR update) and
 * convolution (process) queues.  `version` must have major 1
 * (TANContextImpl.cpp:114-120), else TANX_TAN_UNSUPPORTED_VERSION. */
int32_t tanx_ctx_create(uint64_t version, int32_t device, tanx_ctx **out);
/* TANContext::InitOpenCL(cl_command_queue pConvolutionQueue, cl_command_queue
 * pGeneralQueue) (TANContextImpl.cpp:316-330; the header,
 * TrueAudioNext.h:589-591, names the two parameters the other way round and
 * the implementation wins): like tanx_ctx_create, but Process, group
 * launches and the output chain run on conv_stream and IR updates and
 * transforms on general_stream, hipStream_t of the caller on `device` (the
 * same stream may be given twice; NULL: a stream of the context's own).
 * TANX_INVALID_ARG when a stream belongs to another device.  The context
 * drains the caller's streams when it goes but does not destroy them; they
 * must outlive it. */
int32_t tanx_ctx_create_on_streams(uint64_t version, int32_t device, void *conv_stream, void *general_stream,
                                   tanx_ctx **out);
/* TANContext::Terminate (TANContextImpl.cpp:158-179): releases the caller's
 * handle.  Convolution objects and mixers made on the context hold it too,
 * as the reference's reference-counted objects do, so the context's streams
 * and buffers go when the last of them is destroyed; the handle itself must
 * not be used again. */
int32_t tanx_ctx_destroy(tanx_ctx *ctx);
int32_t tanx_ctx_device(const tanx_ctx *ctx, int32_t *device);
/* The context's process stream (a hipStream_t), on which every Process,
 * IR update (copies and transform) and group launch of its objects runs, in
 * call order; for callers that order their own work or timing events against
 * it (the reference's counterpart is TANContext::GetOpenCLConvQueue,
 * TrueAudioNext.h:596).  A caller that writes inputs or a device IR on its own
 * stream orders this stream after that write. */
void *tanx_ctx_process_stream(tanx_ctx *ctx);
/* The context's general stream (a hipStream_t), on which TANFFT transforms
 * run (TANContext::GetOpenCLGeneralQueue, TrueAudioNext.h:597). */
void *tanx_ctx_update_stream(tanx_ctx *ctx);

/* ---- convolution ----------------------------------------------------------- */

/* TANCreateConvolution + TANConvolution::Init/InitGpu
 * (convolution/ConvolutionImpl.cpp:66-80, 179-305, 1113-1259).  The uniform
 * methods (1, 3, 4, 5, 6) and overlap-add (0) map to the one uniformly
 * partitioned engine; bit 0x8000 selects ProcessFinalize mode.  The
 * non-uniform methods (2, 8, 9; ConvolutionImpl.cpp:201-262, 2380-2643) split
 * the IR into a head of one frame of M blocks, run as uniform partitions of
 * B, and frame partitions of M * B whose multiply-accumulate is spread over
 * the M blocks of a frame; the output is the same convolution with the same
 * B-sample latency, and an IR switch crossfades at the first frame boundary
 * at least M + 1 blocks after the switch would happen in a uniform object
 * (the reference's non-uniform mode also waits for a frame boundary,
 * ConvolutionImpl.cpp:899-905).  B must be a power of two in [64, 1024]
 * (the reference GPU path requires a power of two, ConvolutionImpl.cpp:
 * 1137-1146).  The IR length is rounded up to a power of two as the
 * reference does (ConvolutionImpl.cpp:1230-1235). */
int32_t tanx_conv_create(tanx_ctx *ctx, uint32_t method, uint32_t ir_len,
                         uint32_t block, uint32_t channels, tanx_conv **out);
/* Storage precision of the IR spectra.  FP16 and BF16 rows are scaled by
 * their max |re|,|im| (one FP32 scale per partition row); every product and
 * sum is FP32.  FP32 is the reference's precision and the default. */
enum { TANX_IR_FP32 = 0, TANX_IR_FP16 = 1, TANX_IR_BF16 = 2 };
typedef struct tanx_conv_options {
    uint32_t ir_precision;  /* TANX_IR_* */
    uint32_t nu_multiple;   /* non-uniform methods: blocks per frame (a power
                               of two >= 2, frame <= 4096 samples); 0 = chosen
                               by the library, 1 = uniform partitions */
    uint32_t nu_phase;      /* non-uniform objects: frame phase in blocks,
                               taken modulo M (ignored for uniform objects):
                               the first frame ends at block M-1-(phase % M).
                               Objects of one group with different phases
                               spread their frame work over the ticks. */
    uint32_t pass_blocks;   /* uniform objects: blocks whose partial sums one
                               tail pass computes (1..16; 1 = a tail every
                               block; 0 = chosen by the library: 4 when the
                               IR has 16 or more partitions, else 1).  A
                               pass streams the IR and FDL rows once for K
                               blocks; passes of 5..8 and 9..16 blocks run
                               kernels with more registers per lane, and
                               the heads between passes add up to K + 3
                               partitions themselves. */
    uint32_t pass_phase;    /* uniform objects with multi-block passes: pass
                               phase in blocks, taken modulo pass_blocks K.
                               After block 0, passes run in the launches of
                               blocks b with (b + phase) % K == 0 (a catch-up
                               pass for an IR switch ends on the same grid).
                               Objects of one group with phases 0..K-1 each
                               stream their rows on a different block, so no
                               block carries every object's pass. */
    uint32_t max_tail_chunks; /* uniform objects: at most this many partition
                               chunks per channel in a tail or pass (0 = the
                               library's choice, min(4096 / C, P - 2, 64)).
                               Each chunk keeps its own partial-sum row per
                               block (sized at creation; the next heads read
                               them all), so objects advanced many at a time,
                               whose launches are wide anyway, need few. */
    uint32_t ir_slots;      /* uniform objects: IR spectrum slots, 3 (0 =
                               3) or 2.  With 3, an update issued while a
                               switch is armed goes to the third slot; with
                               2 (a third less HBM per channel) such an
                               update returns TANX_INPUT_FULL and is not
                               made: the armed filter and the current one
                               are both in use until the crossfade block has
                               run (the reference blocks the caller there,
                               ConvolutionImpl.cpp:421-425).  Updates issued
                               before the switch is armed replace the
                               pending one, as with 3. */
    uint32_t reserved[1];   /* zero */
} tanx_conv_options;
/* tanx_conv_create with options (no counterpart in the reference, which
 * stores FP32 spectra only); opt may be NULL. */
int32_t tanx_conv_create_ex(tanx_ctx *ctx, uint32_t method, uint32_t ir_len, uint32_t block,
                            uint32_t channels, const tanx_conv_options *opt, tanx_conv **out);
/* HBM bytes an object of this shape will hold (for admission control; no
 * counterpart in the reference).  Host-memory IR updates add C * rounded
 * ir_len * 4 bytes of staging on first use. */
int32_t tanx_conv_estimate_bytes(uint32_t method, uint32_t ir_len, uint32_t block, uint32_t channels,
                                 const tanx_conv_options *opt, uint64_t *bytes);
/* TANConvolution::Terminate (ConvolutionImpl.cpp:307-355). */
int32_t tanx_conv_destroy(tanx_conv *conv);

/* UpdateResponseTD(float* ppBuffer[], ...) (ConvolutionImpl.cpp:357-382,
 * 407-695): ir[c] holds n <= ir_len host floats, zero padded to the rounded
 * length.  flags may be NULL; channels with STOP_INPUT keep their IR; FLUSH
 * flushes the channel first.  The new IR is crossfaded in two Process calls
 * later, as in the reference when its update thread keeps up. */
int32_t tanx_conv_update_td(tanx_conv *conv, const float *const *ir, size_t n,
                            const uint32_t *flags, uint32_t opflags);
/* Device-memory variant (stands in for UpdateResponseTD(cl_mem[])):
 * channel c's IR at d_ir + c * stride floats, device resident. */
int32_t tanx_conv_update_td_device(tanx_conv *conv, const float *d_ir, size_t stride,
                                   size_t n, const uint32_t *flags, uint32_t opflags);

/* Process(float* in[], float* out[], ...) (ConvolutionImpl.cpp:809-833,
 * 848-1043).  Processes every whole block of n samples per channel, each as
 * one Process step of the reference, and sets processed to n rounded down to
 * a multiple of the block size (TrueAudioNext.h:208-210); n < block
 * processes nothing.  Outputs are complete on return. */
int32_t tanx_conv_process(tanx_conv *conv, const float *const *in, float *const *out,
                          size_t n, const uint32_t *flags, size_t *processed);
/* Device-memory variant (stands in for Process(cl_mem[], cl_mem[])):
 * channel c reads block b at d_in + c * in_stride + b * block and writes
 * d_out + c * out_stride + b * block.  Asynchronous: ordered on the
 * library's process stream; use tanx_conv_synchronize before reading d_out
 * on the host. */
int32_t tanx_conv_process_device(tanx_conv *conv, const float *d_in, size_t in_stride,
                                 float *d_out, size_t out_stride, size_t n,
                                 const uint32_t *flags, size_t *processed);
/* One block for each of n objects in ONE kernel launch (no counterpart in
 * the reference, which runs one engine per object).  Object i reads
 * d_in[i] + c * in_stride[i] and writes d_out[i] + c * out_stride[i] for
 * channel c, with flags[i] (flags or flags[i] may be NULL), exactly as
 * tanx_conv_process_device would for one block; objects keep independent
 * state, so each result equals the object processed alone.  All objects
 * must share the context and IR precision, have B <= 512 and not use
 * ProcessFinalize mode.  Asynchronous on the context's process stream. */
int32_t tanx_group_process_device(uint32_t n, tanx_conv *const *convs, const float *const *d_in,
                                  const size_t *in_stride, float *const *d_out, const size_t *out_stride,
                                  const uint32_t *const *flags);
/* IR updates of several objects of one context in one call (a serving
 * tick's updates): object i is updated as tanx_conv_update_td_device(
 * convs[i], d_irs[i], ir_stride[i], ir_len[i], flags ? flags[i] : NULL,
 * opflags) would update it, in call order, but the IR transforms of all of
 * them run in one launch per partition size and IR type.  No counterpart in
 * the reference, whose updates are per object (ConvolutionImpl.cpp:407-530);
 * for callers that update many objects per audio tick.  Time-domain objects
 * are not accepted (TANX_NOT_SUPPORTED); an object may appear once.  With
 * TANX_OPERATION_FLAG_BLOCK_UNTIL_READY the call returns once every
 * transform has run. */
int32_t tanx_group_update_td_device(uint32_t n, tanx_conv *const *convs, const float *const *d_irs,
                                    const size_t *ir_stride, const size_t *ir_len,
                                    const uint32_t *const *flags, uint32_t opflags);
/* The same group step with no allocation on any call: no counterpart in the
 * reference.  For callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the launch
 * descriptors go to a fixed device scratch array (one per GPU, about 0.6 MB,
 * resolved when the context is created and held until the process ends)
 * through a static host buffer, instead of the staging
 * tanx_group_process_device allocates and grows.  At most 1,024 objects per
 * call (TANX_INVALID_ARG beyond that).  Results are bit-identical to
 * tanx_group_process_device.  Every call on a GPU shares that GPU's scratch
 * array and its contents persist from call to call: calls that overlap on
 * one GPU (concurrent streams or contexts) race on it, so issue them from
 * one stream, one at a time. */
int32_t tanx_group_process_device_noalloc(uint32_t n, tanx_conv *const *convs, const float *const *d_in,
                                          const size_t *in_stride, float *const *d_out, const size_t *out_stride,
                                          const uint32_t *const *flags);
/* ProcessDirect (ConvolutionImpl.cpp:719-762; TrueAudioNext.h ProcessDirect):
 * time-domain objects (TANX_METHOD_TIME_DOMAIN, IR length <= 4096) only,
 * AMF_NOT_SUPPORTED otherwise.  The IR of channel c is ir[c], taps
 * [nz_first_last[2c], nz_first_last[2c+1]) (all taps when NULL); min(n, IR
 * length) samples of every channel are convolved against its history, which
 * advances.  Host buffers; returns with the outputs written. */
int32_t tanx_conv_process_direct(tanx_conv *conv, const float *const *ir, const float *const *in,
                                 float *const *out, size_t n, size_t *processed, const int32_t *nz_first_last);
/* Device form of ProcessDirect: the reference declares a cl_mem overload
 * (TrueAudioNext.h:235-242) whose body returns AMF_NOT_SUPPORTED
 * (ConvolutionImpl.cpp:766-778); here it runs the time-domain kernel on
 * device arrays.  Channel c's IR is ir + c*ir_stride (ir_len taps of
 * tanx_conv_info, or only its nz_first_last range [first, last) is read),
 * its input in + c*in_stride and its output out + c*out_stride (strides in
 * floats; outputs of different channels must not overlap); min(n, ir_len)
 * samples of every channel, enqueued on the context's process stream (the
 * call returns before the outputs are written; tanx_conv_synchronize
 * waits).  nz_first_last: host array of 2*C ints, or NULL. */
int32_t tanx_conv_process_direct_device(tanx_conv *conv, const float *ir, size_t ir_stride, const float *in,
                                        size_t in_stride, float *out, size_t out_stride, size_t n,
                                        size_t *processed, const int32_t *nz_first_last);
/* ProcessFinalize (ConvolutionImpl.cpp:1046-1079): enqueues the tail
 * multiply-accumulate for the next block.  Implicit after each Process unless
 * the object was created with the 0x8000 bit. */
int32_t tanx_conv_finalize(tanx_conv *conv);
/* Flush (ConvolutionImpl.cpp:1261-1325): zero one channel's history. */
int32_t tanx_conv_flush(tanx_conv *conv, uint32_t channel);
/* GetNextFreeChannel (ConvolutionImpl.cpp:1084-1110). */
int32_t tanx_conv_get_next_free_channel(tanx_conv *conv, uint32_t *channel,
                                        const uint32_t *flags);
/* Waits for all work of this object (on the context's process stream). */
int32_t tanx_conv_synchronize(tanx_conv *conv);

typedef struct tanx_conv_info {
    uint32_t ir_len;       /* rounded IR length */
    uint32_t block;        /* B */
    uint32_t channels;     /* C */
    uint32_t partitions;   /* P = ir_len / B */
    uint32_t tail_chunks;  /* partition chunks per channel in the tail kernel */
    uint32_t finalize_mode;
    uint64_t device_bytes; /* HBM held by the object */
    uint64_t ticks;        /* blocks processed */
    uint64_t crossfades;   /* crossfade blocks processed */
    uint32_t ir_precision; /* TANX_IR_* */
    uint32_t nu_multiple;  /* blocks per frame of non-uniform objects, 1 = uniform */
    uint32_t pass_blocks;  /* blocks per tail pass (1 = a tail every block) */
    uint32_t cached_rows;  /* 1: passes read the FDL and IR rows with ordinary loads, because the
                              pass rows of every live object on the context total at most 128 MB
                              (half the Infinity Cache); 0: nontemporal loads */
    uint32_t ir_slots;     /* IR spectrum slots (3, or 2 by tanx_conv_options.ir_slots) */
} tanx_conv_info;
int32_t tanx_conv_get_info(const tanx_conv *conv, tanx_conv_info *info);
/* Sets the partition chunking of the tail kernel (0 = automatic). */
int32_t tanx_conv_set_tail_chunks(tanx_conv *conv, uint32_t chunks);
/* Autotuner: on an object that has processed no block yet, times the fused
 * tick kernel for each power-of-two tail chunking up to the object's maximum
 * and keeps the fastest (the dry runs leave all state unchanged).  Returns
 * the chosen chunk count and its time per launch in microseconds. */
int32_t tanx_conv_autotune(tanx_conv *conv, uint32_t *chunks, double *us_per_block);
/* Device-memory Process runs each block's head and the next block's tail
 * multiply-accumulate in one launch (default 1); 0 launches them separately
 * as the host-memory path does.  Results are identical either way. */
int32_t tanx_conv_set_fused(tanx_conv *conv, uint32_t fused);
/* Device-memory Process over two or more blocks (no channel flags, fused,
 * timing off) replays a captured hipGraph of one tick launch per block,
 * whose arguments come from a descriptor array copied once per batch of up
 * to 64 blocks (default 1); 0 launches each block directly.  Results are
 * identical either way. */
int32_t tanx_conv_set_graph(tanx_conv *conv, uint32_t use_graph);
/* Kernel timing: when enabled, HIP events are recorded on the process
 * stream around every launch.  "head" counts launches that carry the head
 * role (with the next block's tail when fused); "tail" counts separate tail
 * launches, and with multi-block passes (tanx_conv_info.pass_blocks > 1) the
 * fused launches that run a pass.  read_timing synchronizes and returns
 * summed durations and the algorithmic bytes of the timed launches (tail
 * role or pass: FDL + IR rows of partitions 2..P-1 per channel and filter;
 * head role: x_{t-1}, x_t, y, the new and previous FDL rows, partitions 0
 * and 1 of each filter, and with passes the FDL + IR rows of the further
 * partitions the head adds). */
typedef struct tanx_kernel_timing {
    double head_ms;
    double tail_ms;
    uint64_t head_launches;
    uint64_t tail_launches;
    uint64_t head_alg_bytes;
    uint64_t tail_alg_bytes;
    uint64_t blocks;          /* blocks whose heads the timed launches ran */
} tanx_kernel_timing;
int32_t tanx_conv_enable_timing(tanx_conv *conv, uint32_t enable);
int32_t tanx_conv_read_timing(tanx_conv *conv, tanx_kernel_timing *out, uint32_t reset);
/* Timing of the group launches of a context (no counterpart in the
 * reference): when enabled, every group call's kernels stamp HIP events with
 * their own start and end.  read_group_timing synchronizes and returns, in
 * the same struct: head_* = the group tick kernels (duration, launches, and
 * the algorithmic bytes of the roles they ran: each object's head, the
 * further partitions its head added, its tail or multi-block pass and its
 * frame slice), tail_* = the frame kernels of non-uniform objects, blocks =
 * objects advanced. */
int32_t tanx_ctx_enable_group_timing(tanx_ctx *ctx, uint32_t enable);
int32_t tanx_ctx_read_group_timing(tanx_ctx *ctx, tanx_kernel_timing *out, uint32_t reset);
/* Work-queue group launches (no counterpart in the reference, which has no
 * group launches; for batching callers that advance many uneven streams per
 * tick): with enable, tanx_group_process_device runs each of its launches as
 * a persistent kernel of the device's resident workgroups that take the
 * launch's workgroup items from a counter in order (results identical to the
 * plain launches).  The counters are an allocation of the context, made at
 * the first such call; tanx_group_process_device_noalloc always uses the
 * plain launches. */
int32_t tanx_ctx_set_group_queue(tanx_ctx *ctx, uint32_t enable);
/* 1 if the last processed block was a crossfade block. */
int32_t tanx_conv_last_was_crossfade(const tanx_conv *conv);
/* Stream on which process work is ordered (a hipStream_t). */
void *tanx_conv_process_stream(tanx_conv *conv);

/* ---- FFT ------------------------------------------------------------------- */

/* TANFFT::Transform(direction, log2len, channels, float* in[], float* out[])
 * (fft/FFTImpl.cpp:507-575): each channel buffer holds 2^(log2len+1) floats;
 * complex data interleaved; backward transforms scaled by 1/N; planar
 * layouts keep im at N/2+8.  log2len in [1, 19], the reference's largest
 * plan (FFTImpl.h:67; planar: [4, 19]).  Up to 2,048 complex points a
 * channel is one workgroup's LDS transform; above, a four-step transform
 * (column FFTs, FP64 twiddles, row FFTs) through scratch the context keeps
 * (2 x 2^log2len complex per channel, allocated at the first such call). */
int32_t tanx_fft_transform(tanx_ctx *ctx, uint32_t direction, uint32_t log2len,
                           uint32_t channels, const float *const *in, float *const *out);
/* Device-memory variant (stands in for Transform(cl_mem[]) and
 * TransformBatchGPU): channel c at d_in/d_out + c * stride floats. */
int32_t tanx_fft_transform_device(tanx_ctx *ctx, uint32_t direction, uint32_t log2len,
                                  uint32_t channels, const float *d_in, float *d_out,
                                  size_t stride);

const char *tanx_version(void);

/* ---- output chain: mixer and sample converter -------------------------------- */

/* TAN_SAMPLE_TYPE (TrueAudioNext.h:73-77). */
enum { TANX_SAMPLE_FLOAT = 0, TANX_SAMPLE_SHORT = 1 };

typedef struct tanx_mixer tanx_mixer;
/* TANCreateMixer + TANMixer::Init(buffer_size, num_channels)
 * (TrueAudioNext.h:611-635; mixer/MixerImpl.cpp).  The sum is taken in
 * channel order in single precision, as Mixer.cl:26-42 and the CPU Mix do. */
int32_t tanx_mixer_create(tanx_ctx *ctx, size_t buffer_size, int32_t channels, tanx_mixer **out);
int32_t tanx_mixer_destroy(tanx_mixer *mixer);
/* Mix(float* ppBufferInput[], float* ppBufferOutput): host buffers; returns
 * with the output written. */
int32_t tanx_mix(tanx_mixer *mixer, const float *const *in, float *out);
/* Mix(cl_mem pBufferInput[], cl_mem pBufferOutput): device buffers (a host
 * array of device pointers); asynchronous on the context's process stream. */
int32_t tanx_mix_device(tanx_mixer *mixer, const float *const *d_in, float *d_out);
/* Mix(cl_mem pBufferInput, cl_mem pBufferOutput, amf_size inputStride):
 * channel c at d_in + c * in_stride; asynchronous. */
int32_t tanx_mix_device_strided(tanx_mixer *mixer, const float *d_in, size_t in_stride, float *d_out);

/* TANConverter::Convert on device buffers (cl_mem overloads,
 * TrueAudioNext.h:375-398; ConvertGpu, converter/ConverterImpl.cpp:458-556;
 * Converter.cl).  Element i is read at in[i * in_step + in_offset] and
 * written at out[i * out_step + out_offset] (offsets and steps in samples).
 * float -> short: x * 32767 * gain, clamped to [-32768, 32767], converted
 * toward zero; short -> float: x * gain / 32767; same type: copy.
 * gain > 0 and n > 0 (AMF_INVALID_ARG otherwise); other types:
 * AMF_NOT_IMPLEMENTED.  Asynchronous on the process stream unless clipped is
 * given: then it waits and sets *clipped to 1 if a float -> short conversion
 * clamped a sample.  Clipping is reported through *clipped only; the call
 * still returns AMF_OK, as the reference's does. */
int32_t tanx_convert_device(tanx_ctx *ctx, const void *d_in, size_t in_step, size_t in_offset, int32_t in_type,
                            void *d_out, size_t out_step, size_t out_offset, int32_t out_type, size_t n,
                            float gain, int32_t *clipped);
/* TANConverter::Convert on host buffers (TrueAudioNext.h:353-373;
 * ConverterImpl.cpp:217-453): the device path's conversion through staging;
 * returns with the output written.  The reference's own CPU conversion
 * rounds groups of eight samples to nearest and the rest toward zero, and
 * does not report clamping of the groups of eight; this one converts every
 * sample like its GPU kernel. */
int32_t tanx_convert(tanx_ctx *ctx, const void *in, size_t in_step, int32_t in_type, void *out, size_t out_step,
                     int32_t out_type, size_t n, float gain, int32_t *clipped);

/* ---- block DFT on the matrix cores (precision track) ------------------------- */

/* Real-to-complex DFT of n = 2^log2len samples (5 <= log2len <= 11) per
 * channel as a bf16 GEMM on the matrix cores with FP32 accumulation: split =
 * 1 uses one product, split = 3 splits both operands into bf16 high and low
 * parts and accumulates three products (hi hi, hi lo, lo hi).  The DFT
 * tables are computed in FP64.  Channel c reads d_in + c * in_stride and
 * writes bins 0..n/2 as (re, im) pairs at d_out + c * out_stride
 * (out_stride >= n + 2), as TANFFT's R2C direction; d_in must be 16-byte
 * aligned and in_stride a multiple of 4.  No counterpart in the
 * reference (its transforms are FFTs, FFTImpl.cpp:996-1122).  Asynchronous
 * on the context's process stream. */
int32_t tanx_dft_r2c_mfma_device(tanx_ctx *ctx, uint32_t log2len, uint32_t channels, const float *d_in,
                                 size_t in_stride, float *d_out, size_t out_stride, uint32_t split);

/* ---- IR sources: WAV files (host) ------------------------------------------- */

/* Format of a WAV file: ReadWaveFile's outputs (tan/common/wav.cpp:87-98). */
typedef struct tanx_wav_info {
    uint32_t sample_rate;
    uint16_t bits_per_sample; /* 8, 16 or 32 (32: IEEE float) */
    uint16_t channels;
    uint32_t samples;         /* per channel: data bytes * 8 / (bits * channels) */
    uint16_t format_tag;
    uint16_t reserved;
} tanx_wav_info;

/* ReadWaveFile's header walk (wav.cpp:107-171) without the samples.
 * AMF_FILE_NOT_OPEN: no such file; AMF_INVALID_FORMAT: not RIFF/WAVE, no
 * "fmt " or "data" chunk, or bits other than 8, 16, 32. */
int32_t tanx_wav_read_info(const char *path, tanx_wav_info *info);
/* ReadWaveFile (wav.cpp:87-238): de-interleaves the samples into out[c][0 ..
 * min(samples, capacity)) as floats: 8-bit (u - 127) / 256, 16-bit
 * s / 32768, 32-bit the stored float. */
int32_t tanx_wav_read(const char *path, tanx_wav_info *info, float *const *out, size_t capacity);
/* WavContent::ReadWaveFile (wav.cpp:403-505): the interleaved sample bytes
 * (AMF_OUT_OF_RANGE if they exceed capacity_bytes). */
int32_t tanx_wav_read_raw(const char *path, tanx_wav_info *info, void *data, size_t capacity_bytes);
/* WriteWaveFileF (wav.cpp:240-318): planar floats to 8-, 16- or 32-bit
 * samples; 8 and 16 bits clamp to [-1, 1], scale by 127 or 32767 and convert
 * toward zero (8-bit samples are written as signed bytes, as the reference
 * writes them).  Returns AMF_OK on success (the reference's returns false). */
int32_t tanx_wav_write(const char *path, uint32_t sample_rate, uint16_t channels, uint16_t bits, uint32_t samples,
                       const float *const *in);
/* WriteWaveFileS (wav.cpp:320-401): interleaved 16-bit samples as they are;
 * other widths AMF_NOT_IMPLEMENTED (the reference throws). */
int32_t tanx_wav_write_short(const char *path, uint32_t sample_rate, uint16_t channels, uint16_t bits,
                             uint32_t samples, const int16_t *interleaved);

/* ---- IR sources: room responses by image sources (device) ------------------ */

/* generateRoomResponse flags (TrueAudioVR.h:54-56). */
enum { TANX_ROOM_LIMIT_BOUNCES = 0x40, TANX_ROOM_SUPPRESS_DIRECT = 0x80 };
#define TANX_HEAD_FILTER_MAX 64 /* SIMPLEHRTF_FFTLEN */

/* RoomDefinition (TrueAudioVR.h:101-107), same layout: dimensions in meters
 * (x: width, y: height, z: length) and the wall damping factors in [0, 1]
 * (MaterialProperty::damp, in the order mLeft, mRight, mTop, mBottom,
 * mFront, mBack). */
typedef struct tanx_room {
    float width, height, length;
    float damp_left, damp_right, damp_top, damp_bottom, damp_front, damp_back;
} tanx_room;
/* MonoSource (TrueAudioVR.h:69-72): the source position. */
typedef struct tanx_source {
    float x, y, z;
} tanx_source;
/* StereoListener + HeadModel (TrueAudioVR.h:62-67, 76-81), same layout:
 * head position, ear spacing (m), yaw / pitch / roll (degrees) and the head
 * filter pair. */
typedef struct tanx_listener {
    float head_x, head_y, head_z;
    float ear_spacing;
    float yaw, pitch, roll;
    int32_t filter_length;
    float low_pass[TANX_HEAD_FILTER_MAX];
    float high_pass[TANX_HEAD_FILTER_MAX];
} tanx_listener;

/* generateSimpleHeadRelatedTransform (TrueAudioVR.cpp:361-415): fills
 * filter_length (64), low_pass (head shadow, corner 343 / (1.1 * spacing)
 * Hz) and high_pass = impulse - low_pass.  Host only. */
int32_t tanx_room_head_model(float ear_spacing, tanx_listener *listener);
/* generateRoomResponse with GENROOM_USE_GPU_MEM (TrueAudioVR.cpp:659-778;
 * kernels GenerateRoomResponse.cl:126-325, Fill.cl) for `count` rooms,
 * sources and listeners at once: response k's left ear is written to
 * d_out + 2k * out_stride and its right ear to d_out + (2k + 1) *
 * out_stride, response_length floats each (out_stride >= response_length),
 * ready to feed tanx_conv_update_td_device.  Images span -n .. n per axis,
 * n = 1 + response_length / (size / 340 * sample_rate) samples, capped at
 * max_bounces with TANX_ROOM_LIMIT_BOUNCES; TANX_ROOM_SUPPRESS_DIRECT drops
 * the direct path (as the reference's CPU path does).  Contributions are
 * summed in 2^26 fixed point, so the result does not depend on the order of
 * the sums; the sums are read as signed integers (the reference's Fill.cl
 * reads them as unsigned).  AMF_INVALID_ARG for non-positive sizes, damping
 * outside [0, 1] or a room smaller than one sample per axis;
 * AMF_OUT_OF_RANGE for more than 2^31 - 1 images per ear.  Asynchronous on
 * the context's process stream. */
int32_t tanx_room_response_device(tanx_ctx *ctx, uint32_t count, const tanx_room *rooms, const tanx_source *sources,
                                  const tanx_listener *listeners, int32_t sample_rate, int32_t response_length,
                                  uint32_t flags, int32_t max_bounces, float *d_out, size_t out_stride);
/* generateRoomResponse into host memory (the GPU path's read back,
 * TrueAudioVR.cpp:1127-1160): returns with left and right written. */
int32_t tanx_room_response(tanx_ctx *ctx, const tanx_room *room, const tanx_source *source,
                           const tanx_listener *listener, int32_t sample_rate, int32_t response_length,
                           uint32_t flags, int32_t max_bounces, float *left, float *right);

#ifdef __cplusplus
}
#endif
#endif
