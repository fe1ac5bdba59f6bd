// Internal declarations shared by the HIP kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace tanx {

// Spectra are stored "packed": a 2B-point real FFT of a frame of 2B reals
// has B+1 bins whose DC and Nyquist bins are real; element 0 holds
// (DC.re, Nyquist.re) and elements 1..B-1 hold bins 1..B-1.  Every spectrum
// row is therefore exactly B complex floats (B*8 bytes).

// One launch of k_tick runs up to two roles:
//   head (blocks [0, nhead)): channel c = block; R2C of [x_{t-1}, x_t], FDL
//     insert, A_f = sum of tail partials + X_{t-1} H_{f,1} + X_t H_{f,0},
//     C2R, crossfade, output;
//   tail (blocks [nhead, nhead + C * nchunk)): partial sums over partitions
//     [2 + j*pc, 2 + (j+1)*pc) of X_{t+1-i} H_i for the NEXT block.
// The two roles touch disjoint data, so they may share a launch.
// IR spectrum storage: FP32 complex, or FP16 / BF16 complex scaled per
// partition row (value = stored * S[c * P + i], S = the row's max |re|,|im|),
// always accumulated in FP32.
enum IrType { IR_F32 = 0, IR_F16 = 1, IR_BF16 = 2 };
inline int ir_elem_bytes(int t) { return t == IR_F32 ? 8 : 4; }

struct TickArgs {
    int B;                 // block size (group launches dispatch on it)
    int C;
    int P;
    int irtype;            // IrType of Ha/Hb/tHa/tHb
    const uint8_t *run;    // [C] 1 = channel running
    const float2 *twM;     // W_B^t, t < B
    const float2 *twN;     // W_2B^t, t < 2B
    float2 *fdl;           // [C][P][B] ring of input spectra

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
};

struct IrArgs {
    const float *ir;       // channel ch IR at ir + ch * stride, n valid samples
    long long stride;
    long long n;
    const int *chan;       // list of channels to transform (blockIdx.y indexes it)
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
};

// Launchers (return hipError_t).  B is the block size (power of two).
hipError_t launch_tick(int B, int tail_nf, const TickArgs &a, hipStream_t s);
// One launch for several objects: block b belongs to the object o with
// starts[o] <= b < starts[o + 1] and runs its roles from d_descs[o].  Every
// object must use irtype and a block size in {64, 128, 256, 512}.
hipError_t launch_group_tick(int irtype, const TickArgs *d_descs, const int *d_starts, int nobj, unsigned blocks,
                             hipStream_t s);
constexpr int kGroupMaxB = 512;
// k_tick reading its arguments from d_descs[k] (graph replay)
hipError_t launch_tick_desc(int B, int irtype, const TickArgs *d_descs, int k, unsigned blocks, hipStream_t s);
hipError_t launch_ir_transform(int B, int nchan, const IrArgs &a, hipStream_t s);
hipError_t launch_fft(int channels, const FftArgs &a, hipStream_t s);
hipError_t launch_zero(void *p, size_t bytes, hipStream_t s);

// Twiddle tables: tw + (2^l - 1) holds W_{2^l}^t = exp(-2 pi i t / 2^l), t < 2^l,
// for l = 0..kMaxTwLog2.
constexpr int kMaxTwLog2 = 13;
inline const float2 *tw_table(const float2 *tw, int log2) { return tw + ((1 << log2) - 1); }

}  // namespace tanx
