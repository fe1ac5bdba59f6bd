/*
 * tan_oracle.h -- CPU oracle for the TAN time-varying convolution path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code.  It is used by tests/ (as the checker), by __graft_entry__.smoke()
 * and by the cpu_baseline leg of bench.py (as the timed CPU path).
 *
 * Contents
 *   1. FP64 references: O(N^2) DFT, direct convolution.  These are the
 *      mathematical definitions; they pin everything else.
 *   2. A restatement of the reference CPU FFT interface (TANFFT::Transform,
 *      tan/tanlibrary/src/TrueAudioNext/fft/FFTImpl.cpp:507-575, 876-1122):
 *      same direction codes, buffer layouts and 1/N backward scaling.  The
 *      reference delegates the arithmetic to FFTW 3.3.5 (dlopen of
 *      libfftw3f.so, FFTImpl.cpp:298-346), which is not vendored in the
 *      reference and not installed here, so the transform itself is this
 *      file's own radix-2 FFT.  Parity of this layer is pinned by the FP64
 *      DFT (closed form), not by FFTW's rounding (FFTW's bits: parity
 *      unpinned).
 *   3. A restatement of the CPU uniform-partitioned convolution
 *      (TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_UNIFORM):
 *      ovlNUPProcessCPU / ovlNUPProcessTail (ConvolutionImpl.cpp:2380-2643),
 *      PlanarComplexMultiplyAccumulate (MathImpl.cpp:418-553), the Process
 *      state machine and IR slot rotation (ConvolutionImpl.cpp:848-1043),
 *      UpdateResponseTD + update thread (ConvolutionImpl.cpp:407-557,
 *      2004-2160), Crossfade (ConvolutionImpl.cpp:1353-1411), channel flags
 *      (ConvolutionImpl.cpp:2868-2931) and Flush (ConvolutionImpl.cpp:1261-1325).
 *      Pinned by the FP64 direct convolution and by closed-form crossfade
 *      weights.
 *   4. The device I/O chain (mixer, sample converter).
 *   5. Room impulse responses by image sources (tan_oracle_room.c), pinned
 *      by a textbook image-source model built from reflection geometry and
 *      by closed forms (direct path alone, one-axis echo trains, the head
 *      filters against numpy's inverse FFT).
 */
#ifndef TAN_ORACLE_H
#define TAN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1. FP64 references ------------------------------------------------ */

/* X[k] = sum_n x[n] exp(sign * 2 pi i n k / n_len), sign = -1 forward,
 * +1 inverse (unscaled).  Interleaved complex (re, im) doubles. */
void orc_dft_f64(int n_len, const double *in, double *out, int inverse);

/* y[n] = sum_{k=0}^{nh-1} h[k] x[n-k], n in [0, ny), x[m] = 0 outside
 * [0, nx).  Accumulated in double.  OpenMP over n. */
void orc_direct_conv_f64(const float *x, long nx, const float *h, long nh,
                         double *y, long ny);

/* ---- 2. TANFFT::Transform restatement ----------------------------------- */

/* direction codes as TAN_FFT_TRANSFORM_DIRECTION (TrueAudioNext.h:493-501) */
enum {
    ORC_FFT_FORWARD = 0,
    ORC_FFT_BACKWARD = 1,
    ORC_FFT_R2C_FORWARD = 2,
    ORC_FFT_C2R_BACKWARD = 3,
    ORC_FFT_R2C_PLANAR_FORWARD = 4,
    ORC_FFT_C2R_PLANAR_BACKWARD = 5
};

/* Same buffer conventions as TANFFT::Transform: each channel buffer holds
 * 2^(log2len+1) floats.  Complex directions: interleaved N complex; backward
 * scaled by 1/N.  R2C: N reals in, N/2+1 interleaved complex out.  Planar:
 * re at [0, N/2], im at [N/2+8, N+8].  C2R outputs N reals scaled by 1/N.
 * Returns 0 (AMF_OK) or 4 (AMF_INVALID_ARG). */
int orc_fft_transform(int direction, int log2len, int channels,
                      float *const *in, float *const *out);

/* ---- 3. Uniform-partitioned convolution restatement ---------------------- */

enum {
    ORC_XFADE_INTENDED = 0,  /* exact output-switched crossfade, one history */
    ORC_XFADE_REFERENCE = 1  /* the reference's crossfade block as written:
                                two ProcessInternal calls, two FDL advances
                                (ConvolutionImpl.cpp:976-991) */
};

typedef struct orc_conv orc_conv;

/* InitCpu(FFT_PARTITIONED_UNIFORM, L, B, C) (ConvolutionImpl.cpp:179-305,
 * 1113-1259).  B must be a power of two.  finalize_mode mirrors the
 * TAN_CONVOLUTION_METHOD_USE_PROCESS_FINALIZE bit.  NULL on bad args. */
orc_conv *orc_conv_create(int L, int B, int C, int xfade_mode,
                          int finalize_mode);
void orc_conv_destroy(orc_conv *c);

/* AMF_RESULT-valued, same argument meaning as TANConvolution. */
int orc_conv_update_td(orc_conv *c, const float *const *ir, size_t n,
                       const uint32_t *flags, uint32_t opflags);
int orc_conv_process(orc_conv *c, const float *const *in, float *const *out,
                     size_t n, const uint32_t *flags, size_t *processed);
int orc_conv_finalize(orc_conv *c);
int orc_conv_get_next_free_channel(orc_conv *c, uint32_t *idx,
                                   const uint32_t *flags);
/* 1 if the last orc_conv_process call was a crossfade block. */
int orc_conv_last_was_crossfade(const orc_conv *c);

/* Planted-mistake switch for the oracle's own self-tests: 0 = none,
 * 1 = conjugated IR in the CMAC, 2 = last partition dropped from the tail,
 * 3 = FDL ring walked the wrong way, 4 = C2R missing the 1/N scale,
 * 5 = crossfade weights swapped. */
void orc_set_mutation(int m);

/* OpenMP thread count used by the restated CMAC/FFT loops
 * (the reference sets omp_get_num_procs()/4: FFTImpl.cpp:245-250,
 * MathImpl.cpp:154-158).  Returns the count now in effect. */
int orc_set_threads(int n);
int orc_num_procs(void);

/* CPU baseline timing: runs the restated CPU path for nblocks blocks of a
 * (C, L, B) instance with one IR (seeded decaying noise) and white-noise
 * input, after `warm` untimed blocks; fills per-block wall times (seconds,
 * Process + finalize) into times[nblocks].  Returns total seconds. */
double orc_bench_upols(int C, int L, int B, int nblocks, int warm,
                       uint64_t seed, double *times);

/* 4. The device I/O chain of the reference GPU path.
 *
 * TANMixer::Mix: out[s] = in_0[s] + in_1[s] + ... in channel order, single
 * precision, each partial sum rounded (mixer/Mixer.cl:26-42;
 * MixerImpl.cpp Mix(float**, float*) adds in the same order). */
void orc_mix(const float *const *in, int channels, size_t n, float *out);

/* TANConverter on the GPU (converter/Converter.cl, ConvertGpu at
 * ConverterImpl.cpp:458-556).  Types: 0 = float, 1 = short
 * (TAN_SAMPLE_TYPE, TrueAudioNext.h:73-77).  float -> short:
 * f = x * (32767 * gain), clamped to [-32768, 32767] (clamping sets
 * *clipped), then converted toward zero (OpenCL convert_short);
 * short -> float: x * (gain / 32767); same type: copy.  Element i is read
 * at in[i * in_step + in_offset] and written at out[i * out_step +
 * out_offset].  Returns 0, or -1 for an unknown type. */
int orc_convert(const void *in, size_t in_step, size_t in_offset, int in_type, void *out, size_t out_step,
                size_t out_offset, int out_type, size_t n, float gain, int *clipped);

/* 5. Room impulse responses (tan_oracle_room.c): TrueAudioVR's GPU path
 * (generateRoomResponse host set-up, TrueAudioVR.cpp:659-778; kernels
 * GenerateRoomResponse.cl:204-325 and Fill.cl) restated in single
 * precision without contraction, and the head filters of
 * generateSimpleHeadRelatedTransform (TrueAudioVR.cpp:361-415).  The
 * structs have the layouts of RoomDefinition, MonoSource and
 * StereoListener (TrueAudioVR.h:62-107).  Return values are AMF_RESULT
 * codes (0 ok, 4 invalid argument, 5 out of range, 6 out of memory). */
#define ORC_HEAD_FILTER 64
#define ORC_ROOM_LIMIT_BOUNCES 0x40
#define ORC_ROOM_SUPPRESS_DIRECT 0x80
typedef struct orc_room {
    float width, height, length;
    float damp_left, damp_right, damp_top, damp_bottom, damp_front, damp_back;
} orc_room;
typedef struct orc_source {
    float x, y, z;
} orc_source;
typedef struct orc_listener {
    float head_x, head_y, head_z;
    float ear_spacing;
    float yaw, pitch, roll;
    int32_t filter_length;
    float low_pass[ORC_HEAD_FILTER];
    float high_pass[ORC_HEAD_FILTER];
} orc_listener;
/* Kernel arguments of one ear (GenerateRoomResponse's parameter list). */
typedef struct orc_room_ear {
    float src[3], head[3], ev[3], ev_norm, dim[3];
    float damp[6]; /* right, left, top, bottom, front, back */
    float max_gain, d_min;
    int fs, len, hrtf_len, filt_len;
    int n[3];
    int suppress_direct;
    float lp[ORC_HEAD_FILTER], hp[ORC_HEAD_FILTER];
} orc_room_ear;
int orc_room_ear_params(const orc_room *room, const orc_source *source, const orc_listener *ears, int fs, int len,
                        unsigned flags, int max_bounces, orc_room_ear out[2], long long *images);
/* One image: 0 = no contribution, 1 = one tap at *index of value
 * (int)*scale, 2 = *taps head-filtered taps from *index of value
 * (int)(*scale * (lp[i] + *hf * hp[i])). */
int orc_room_image(const orc_room_ear *ear, int ix, int iy, int iz, int *index, int *taps, float *scale, float *hf);
int orc_room_response(const orc_room *room, const orc_source *source, const orc_listener *listener, int fs,
                      int len, unsigned flags, int max_bounces, float *left, float *right);
int orc_room_head_model(float ear_spacing, orc_listener *listener);

#ifdef __cplusplus
}
#endif
#endif
