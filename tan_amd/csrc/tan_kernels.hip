// gfx950 kernels of the uniformly partitioned FFT convolution path.
//
// Method: overlap-save with a frequency-domain delay line (FDL).
//   X_t    = FFT_2B([x_{t-1}, x_t])                       (head)
//   A_f(t) = sum_{i=0}^{P-1} X_{t-i} H_{f,i}              (tail: i >= 1, head: i = 0)
//   y_f(t) = IFFT_2B(A_f(t))[B .. 2B)                     (head)
// with H_{f,i} = FFT_2B([h_f[iB .. iB+B), 0_B]).  The reference CPU path
// (ovlNUPProcessCPU/ovlNUPProcessTail, ConvolutionImpl.cpp:2380-2643) is the
// overlap-add form of the same sum; overlap-save keeps no filter-dependent
// output state, so switching IRs (Process crossfade, ConvolutionImpl.cpp:
// 893-991) evaluates both filters exactly on one input history.
//
// FFTs: the 2B-point real transforms run as B-point complex Stockham
// radix-4 (+ one radix-2) FFTs staged in LDS, with the standard real/complex
// split post- and pre-processing.  Twiddles come from tables computed in
// FP64 on the host and rounded to FP32.
#include "tan_internal.h"

#include <hip/hip_ext.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <map>
#include <mutex>

namespace tanx {

// Loads and stores through the global address space.  A pointer read from
// device memory (a group descriptor) is generic to the compiler, and plain
// accesses through it become flat instructions: 64-bit per-lane addresses
// (two VGPRs each) and waits shared with LDS.  Through these helpers they
// are global instructions with a scalar base and a 32-bit lane offset.
template <class T>
using gmem_t = __attribute__((address_space(1))) T;
template <class T>
__device__ __forceinline__ T gld(const T *p)
{
    return *(const gmem_t<T> *)p;
}
template <class T>
__device__ __forceinline__ void gst(T *p, T v)
{
    *(gmem_t<T> *)p = v;
}
// float2 / float4 go through native vectors: the HIP vector classes copy
// through a generic reference, which would make the access flat again
typedef float v2f __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float2 gld(const float2 *p)
{
    const v2f v = *(const gmem_t<v2f> *)p;
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float4 gld(const float4 *p)
{
    const v4f v = *(const gmem_t<v4f> *)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst(float2 *p, float2 v)
{
    v2f w;
    w.x = v.x;
    w.y = v.y;
    *(gmem_t<v2f> *)p = w;
}
__device__ __forceinline__ void gst(float4 *p, float4 v)
{
    v4f w;
    w.x = v.x;
    w.y = v.y;
    w.z = v.z;
    w.w = v.w;
    *(gmem_t<v4f> *)p = w;
}

// Products written with explicit fused multiply-adds: the rounding is then
// the same in every kernel that inlines them (a contracted a*b + c*d could
// fuse either product, and the choice can differ between instantiations).
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)  // a * conj(b)
{
    return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -(a.x * b.y)));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Stockham autosort FFTs of nb independent M-point sequences held in LDS:
// sequence j at a + j*M, with scratch b + j*M.  Returns the base (a or b)
// that holds the results (natural order).  Radix-4 stages with a trailing
// radix-2 stage when log2(M) is odd.  Each stage: butterfly q of a sequence
// reads a[q + m*M/R], m < R, multiplies by W_{pR}^{m*k} (k = q mod p), and
// writes b[(q - k)*R + k + m*p]; the butterflies of all nb sequences are
// spread over the block's threads, so nb sequences take the stages (and
// barriers) of one.  INV selects the unscaled inverse.  All threads of the
// block must call it; with NT = 64 the calling wave runs the sequences alone
// (its own buffers, a wave barrier between stages) and only its 64 lanes
// call it.
constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }

// Orders one wave's LDS accesses: its lanes' stores before the barrier are
// seen by its lanes' loads after it (a wave's LDS operations complete in
// issue order; the fences keep the compiler from moving accesses across).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Barrier of the threads that share a piece of work: the wave (NT = 64) or
// the workgroup.
template <int NT>
__device__ __forceinline__ void group_sync()
{
    if constexpr (NT == 64) wave_sync();
    else __syncthreads();
}

template <int M, bool INV, int NT = 0>
__device__ float2 *fft_lds_n(float2 *a, float2 *b, int nb, const float2 *__restrict__ tw)
{
    constexpr int LQ = ilog2(M / 4), LH = ilog2(M / 2);  // butterflies per sequence: radix 4, radix 2
    const int tid = NT == 64 ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
    const int nt = NT == 64 ? 64 : (int)blockDim.x;
    int p = 1;
#pragma unroll 1
    for (; p * 4 <= M; p *= 4) {
        const int tstride = M / (4 * p);
        for (int e = tid; e < (nb << LQ); e += nt) {
            const int q = e & (M / 4 - 1);
            const float2 *aj = a + (e >> LQ) * M;
            float2 *bj = b + (e >> LQ) * M;
            const int k = q & (p - 1);
            float2 x0 = aj[q], x1 = aj[q + M / 4], x2 = aj[q + M / 2], x3 = aj[q + 3 * M / 4];
            if (p > 1) {
                const int t = k * tstride;
                float2 w1 = tw[t], w2 = tw[2 * t], w3 = tw[3 * t];
                if (INV) { w1.y = -w1.y; w2.y = -w2.y; w3.y = -w3.y; }
                x1 = cmul(x1, w1);
                x2 = cmul(x2, w2);
                x3 = cmul(x3, w3);
            }
            const float2 s02 = cadd(x0, x2), d02 = csub(x0, x2);
            const float2 s13 = cadd(x1, x3), d13 = csub(x1, x3);
            // forward: W_4 = -i ; inverse: +i
            const float2 jd13 = INV ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
            const int j = (q - k) * 4 + k;
            bj[j] = cadd(s02, s13);
            bj[j + p] = cadd(d02, jd13);
            bj[j + 2 * p] = csub(s02, s13);
            bj[j + 3 * p] = csub(d02, jd13);
        }
        group_sync<NT>();
        float2 *t = a; a = b; b = t;
    }
    if (p < M) {  // radix-2, p = M/2
        for (int e = tid; e < (nb << LH); e += nt) {
            const int q = e & (M / 2 - 1);
            const float2 *aj = a + (e >> LH) * M;
            float2 *bj = b + (e >> LH) * M;
            float2 x0 = aj[q], x1 = aj[q + M / 2];
            float2 w = tw[q];
            if (INV) w.y = -w.y;
            x1 = cmul(x1, w);
            bj[q] = cadd(x0, x1);
            bj[q + M / 2] = csub(x0, x1);
        }
        group_sync<NT>();
        a = b;
    }
    return a;
}

// One M-point FFT (see fft_lds_n).
template <int M, bool INV, int NT = 0>
__device__ __forceinline__ float2 *fft_lds(float2 *a, float2 *b, const float2 *__restrict__ tw)
{
    return fft_lds_n<M, INV, NT>(a, b, 1, tw);
}

// Real-FFT post-processing: Z = FFT_M(z), z[n] = x[2n] + i x[2n+1].
// Returns packed bin k of the 2M-point real FFT (k = 0: (X[0], X[M])); w is
// the 2M-point twiddle of bin k (unused for k = 0).
template <int M>
__device__ __forceinline__ float2 r2c_post_w(const float2 *Z, int k, float2 w)
{
    const float2 z0 = Z[k];
    if (k == 0) return make_float2(z0.x + z0.y, z0.x - z0.y);
    const float2 zc = Z[M - k];                       // conj taken below
    const float er = 0.5f * (z0.x + zc.x), ei = 0.5f * (z0.y - zc.y);
    const float dr = z0.x - zc.x, di = z0.y + zc.y;   // Z[k] - conj(Z[M-k])
    const float orr = 0.5f * di, oi = -0.5f * dr;     // (Z - conj)/(2i)
    return make_float2(er + fmaf(w.x, orr, -(w.y * oi)), ei + fmaf(w.x, oi, w.y * orr));
}

template <int M>
__device__ __forceinline__ float2 r2c_post(const float2 *Z, int k, const float2 *__restrict__ twN)
{
    return k == 0 ? r2c_post_w<M>(Z, 0, make_float2(1.f, 0.f)) : r2c_post_w<M>(Z, k, twN[k]);
}

// Inverse pre-processing from a packed spectrum A (element 0 = (DC, Nyquist)):
// returns Z''[k] = 2 (E[k] + i O[k]) so that IFFT_M(Z'')/(2M) interleaves the
// even/odd output samples; w as in r2c_post_w.
template <int M>
__device__ __forceinline__ float2 c2r_pre_w(const float2 *A, int k, float2 w)
{
    const float2 a = A[k];
    if (k == 0) return make_float2(a.x + a.y, a.x - a.y);
    const float2 c = A[M - k];                        // conj(c) used
    const float er = a.x + c.x, ei = a.y - c.y;
    const float dr = a.x - c.x, di = a.y + c.y;       // X[k] - conj(X[M-k])
    // multiply by conj(w)
    const float orr = fmaf(dr, w.x, di * w.y), oi = fmaf(di, w.x, -(dr * w.y));
    return make_float2(er - oi, ei + orr);
}

template <int M>
__device__ __forceinline__ float2 c2r_pre(const float2 *A, int k, const float2 *__restrict__ twN)
{
    return k == 0 ? c2r_pre_w<M>(A, 0, make_float2(1.f, 0.f)) : c2r_pre_w<M>(A, k, twN[k]);
}

// ---------------------------------------------------------------------------
// Tick kernel: head and tail roles (see TickArgs).  256 threads per block.
// ---------------------------------------------------------------------------
constexpr int kTickThreads = 256;
constexpr int kHeadChunkBatch = 16;  // partial-sum rows loaded together by the head
constexpr int kHeadPreload = 4;      // near partitions a head loads together

// float4 of LDS a tail or pass role reduces its lane groups' sums through:
// [filter (<= 2)][lane group G][row float4 H] per round, G * H = 256 * V
template <int M>
constexpr int tail_red_float4() { return 2 * kTickThreads * ((M / 2 > 256) ? (M / 2) / 256 : 1); }

// Streamed rows (IR spectra, FDL) are read once per block: nontemporal loads
// keep them from displacing reused data in the caches.  A multi-block pass
// over an object whose rows fit in the Infinity Cache reads them with
// ordinary loads instead (ld_row<false>), so they stay resident from one
// launch to the next.
__device__ __forceinline__ float4 ld_stream(const float4 *p)
{
    const v4f v = __builtin_nontemporal_load((const gmem_t<v4f> *)p);
    return make_float4(v.x, v.y, v.z, v.w);
}
template <bool NT>
__device__ __forceinline__ float4 ld_row(const float4 *p)
{
    if constexpr (NT) return ld_stream(p);
    return gld(p);
}

typedef unsigned int v2u __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float half_lo(unsigned int w) { return __half2float(__ushort_as_half((unsigned short)(w & 0xffffu))); }
__device__ __forceinline__ float half_hi(unsigned int w) { return __half2float(__ushort_as_half((unsigned short)(w >> 16))); }
__device__ __forceinline__ float bf_lo(unsigned int w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(unsigned int w) { return __uint_as_float(w & 0xffff0000u); }

// Access to one IR spectrum row (B complex) of a storage type.
template <int HT>
struct IrRow;

template <>
struct IrRow<IR_F32> {
    static constexpr int kBytes = 8;
    __device__ static float2 get1(const void *row, int k) { return gld(reinterpret_cast<const float2 *>(row) + k); }
    // complex 2q and 2q+1, streamed (getc: ordinary load)
    __device__ static float4 get2(const void *row, int q) { return ld_stream(reinterpret_cast<const float4 *>(row) + q); }
    __device__ static float4 getc(const void *row, int q) { return gld(reinterpret_cast<const float4 *>(row) + q); }
};

template <>
struct IrRow<IR_F16> {
    static constexpr int kBytes = 4;
    __device__ static float2 get1(const void *row, int k)
    {
        const unsigned int w = gld(reinterpret_cast<const unsigned int *>(row) + k);
        return make_float2(half_lo(w), half_hi(w));
    }
    __device__ static float4 get2(const void *row, int q)
    {
        const v2u w = __builtin_nontemporal_load((const gmem_t<v2u> *)row + q);
        return make_float4(half_lo(w.x), half_hi(w.x), half_lo(w.y), half_hi(w.y));
    }
    __device__ static float4 getc(const void *row, int q)
    {
        const v2u w = gld(reinterpret_cast<const v2u *>(row) + q);
        return make_float4(half_lo(w.x), half_hi(w.x), half_lo(w.y), half_hi(w.y));
    }
};

template <>
struct IrRow<IR_BF16> {
    static constexpr int kBytes = 4;
    __device__ static float2 get1(const void *row, int k)
    {
        const unsigned int w = gld(reinterpret_cast<const unsigned int *>(row) + k);
        return make_float2(bf_lo(w), bf_hi(w));
    }
    __device__ static float4 get2(const void *row, int q)
    {
        const v2u w = __builtin_nontemporal_load((const gmem_t<v2u> *)row + q);
        return make_float4(bf_lo(w.x), bf_hi(w.x), bf_lo(w.y), bf_hi(w.y));
    }
    __device__ static float4 getc(const void *row, int q)
    {
        const v2u w = gld(reinterpret_cast<const v2u *>(row) + q);
        return make_float4(bf_lo(w.x), bf_hi(w.x), bf_lo(w.y), bf_hi(w.y));
    }
};

// base address of row (c, i) of a slot
template <int HT>
__device__ __forceinline__ const void *ir_row(const void *base, int c, int P, int i, int M)
{
    return reinterpret_cast<const char *>(base) + ((size_t)c * P + i) * (size_t)M * IrRow<HT>::kBytes;
}

template <int HT>
__device__ __forceinline__ float ir_scale(const float *S, int c, int P, int i)
{
    if constexpr (HT == IR_F32) {
        return 1.0f;
    } else {
        return gld(S + (size_t)c * P + i);
    }
}

__device__ __forceinline__ float4 scale4(float4 v, float s) { return make_float4(v.x * s, v.y * s, v.z * s, v.w * s); }

__device__ __forceinline__ void mac2(float2 &acc, float2 x, float2 h, bool dc)
{
    // packed DC/Nyquist element: two real products; otherwise complex
    acc.x = fmaf(x.x, h.x, acc.x);
    if (dc) {
        acc.y = fmaf(x.y, h.y, acc.y);
    } else {
        acc.x = fmaf(-x.y, h.y, acc.x);
        acc.y = fmaf(x.x, h.y, acc.y);
        acc.y = fmaf(x.y, h.x, acc.y);
    }
}

__device__ __forceinline__ void mac4(float4 &acc, float4 x, float4 h, float s, bool dc)
{
    // complex 0 (packed DC/Nyquist when dc; s = 0 then)
    acc.x = fmaf(x.x, h.x, acc.x);
    acc.x = fmaf(-s * x.y, h.y, acc.x);
    acc.y = fmaf(s * x.x, h.y, acc.y);
    acc.y = fmaf(x.y, dc ? h.y : h.x, acc.y);
    // complex 1
    acc.z = fmaf(x.z, h.z, acc.z);
    acc.z = fmaf(-x.w, h.w, acc.z);
    acc.w = fmaf(x.z, h.w, acc.w);
    acc.w = fmaf(x.w, h.z, acc.w);
}

// LDS of one head: the FFT's two buffers and the M-point twiddles.  The
// bins a thread owns (k = tid + q * NT) stay in its registers between the
// transforms: the forward spectrum, the 2M-point twiddles and the output.
template <int M>
struct HeadSmem {
    float2 s0[M], s1[M], tw[M];
};

// Head of channel c, run by NT threads: one wave (NT = 64, with its own
// HeadSmem and wave barriers) or the whole workgroup (NT = kTickThreads).
template <int M, int HT, int NT>
__device__ void head_role(const TickArgs &a, int c, HeadSmem<M> &sm)
{
    constexpr int KPT = M >= NT ? M / NT : 1;  // bins per thread
    constexpr int ZP = (M / 2 + NT - 1) / NT;  // output sample pairs per thread
    constexpr float kScale = 1.0f / (2.0f * (float)M);
    // rows per memory round trip: partial sums (CB) and near partitions
    // (NB); a wave's lane holds KPT bins of each row
    constexpr int CB = NT == 64 ? 2 : kHeadChunkBatch;
    constexpr int NB = NT == 64 ? 2 : kHeadPreload;
    static_assert(M % NT == 0 || NT > M, "bins per thread");
    const int tid = NT == 64 ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
    const int nf = a.xfade ? 2 : 1;
    // the head is the launch's longest chain: its waves issue before the
    // tail's on a shared compute unit
    __builtin_amdgcn_s_setprio(3);

    // ---- issue every global load first: frame, twiddles, near rows, partials
    const float *xp = a.xprev + (size_t)c * M;
    const float *xi = a.xin + (size_t)c * a.in_stride;
    for (int n = tid; n < M; n += NT) {
        const float *src = n < M / 2 ? xp + 2 * n : xi + 2 * (n - M / 2);
        sm.s0[n] = make_float2(gld(src), gld(src + 1));
        sm.tw[n] = gld(a.twM + n);
    }
    float2 wn[KPT];
#pragma unroll
    for (int q = 0; q < KPT; q++) wn[q] = gld(a.twN + min(tid + q * NT, M - 1));
    const int dmax = max(a.dA, nf == 2 ? a.dB : 1);  // 1: partition 1 only
    float2 acc[2][KPT], h0[2][KPT];
    // bin of slot q, clamped for the loads of threads past M (NT > M:
    // their values are never used)
    auto bin = [&](int q) { return min(tid + q * NT, M - 1); };
#pragma unroll
    for (int f = 0; f < 2; f++) {
        if (f >= nf) continue;
        const void *Hbase = f == 0 ? a.Ha : a.Hb;
        const float *Sbase = f == 0 ? a.Sa : a.Sb;
        const void *H0row = ir_row<HT>(Hbase, c, a.P, 0, M);
        const float s0 = ir_scale<HT>(Sbase, c, a.P, 0);
        const float2 *T = (f == 0 ? a.Ta : a.Tb) + (size_t)c * a.tchunk * M;
        // partition 0 and the partial sums: CB rows per memory round trip
        // (clamped row indices keep every load in range; rows past tchunk
        // add zero)
        float2 s[KPT];
#pragma unroll
        for (int q = 0; q < KPT; q++) {
            const float2 h0v = IrRow<HT>::get1(H0row, bin(q));
            h0[f][q] = make_float2(h0v.x * s0, h0v.y * s0);
            s[q] = make_float2(0.f, 0.f);
        }
        for (int j = 0; j < a.tchunk; j += CB) {
            float2 t[CB][KPT];
#pragma unroll
            for (int u = 0; u < CB; u++)
#pragma unroll
                for (int q = 0; q < KPT; q++) t[u][q] = gld(T + (size_t)min(j + u, a.tchunk - 1) * M + bin(q));
#pragma unroll
            for (int u = 0; u < CB; u++)
#pragma unroll
                for (int q = 0; q < KPT; q++) {
                    const bool in = j + u < a.tchunk;
                    s[q].x += in ? t[u][q].x : 0.f;
                    s[q].y += in ? t[u][q].y : 0.f;
                }
        }
        // partitions d..1, descending (the order heads_role keeps): the
        // blocks since the launch whose pass computed the partials, NB rows
        // per memory round trip
        const int d = a.P > 1 ? (dmax == 1 ? 1 : min(f == 0 ? a.dA : a.dB, a.P - 1)) : 0;
        for (int jb = d; jb >= 1; jb -= NB) {
            float2 fr[NB][KPT], hr[NB][KPT];
            float sj[NB];
#pragma unroll
            for (int u = 0; u < NB; u++) {
                const int jn = max(jb - u, 1);
                const float2 *Fj = a.fdl + ((size_t)c * a.P + ((a.slot - jn) & (a.P - 1))) * M;
                const void *Hj = ir_row<HT>(Hbase, c, a.P, jn, M);
                sj[u] = ir_scale<HT>(Sbase, c, a.P, jn);
#pragma unroll
                for (int q = 0; q < KPT; q++) {
                    fr[u][q] = gld(Fj + bin(q));
                    hr[u][q] = IrRow<HT>::get1(Hj, bin(q));
                }
            }
#pragma unroll
            for (int u = 0; u < NB; u++) {
                if (jb - u >= 1) {
#pragma unroll
                    for (int q = 0; q < KPT; q++) {
                        const int k = tid + q * NT;
                        mac2(s[q], fr[u][q], make_float2(hr[u][q].x * sj[u], hr[u][q].y * sj[u]), k == 0);
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < KPT; q++) acc[f][q] = s[q];
    }
    // non-uniform objects: the frame outputs for this block (read early)
    float2 za[ZP], zb[ZP];
#pragma unroll
    for (int q = 0; q < ZP; q++) {
        za[q] = make_float2(0.f, 0.f);
        zb[q] = make_float2(0.f, 0.f);
        const int n = tid + q * NT;
        if (n < M / 2) {
            if (a.zA) za[q] = gld(reinterpret_cast<const float2 *>(a.zA + (size_t)c * a.N + a.zoff) + n);
            if (a.zB) zb[q] = gld(reinterpret_cast<const float2 *>(a.zB + (size_t)c * a.N + a.zoff) + n);
        }
    }
    group_sync<NT>();
    if (tid < M) {
        // this block becomes the previous one of the next frame
        float *xpw = a.xprev + (size_t)c * M;
        for (int j = tid; j < M; j += NT) gst(xpw + j, gld(xi + j));
        if (a.nu) {
            float *xr = a.xring + ((size_t)c * 2 * a.M + a.xslot) * M;
            for (int j = tid; j < M; j += NT) gst(xr + j, gld(xi + j));
        }
    }

    const float2 *Z = fft_lds<M, false, NT>(sm.s0, sm.s1, sm.tw);
    float2 *row = a.fdl + ((size_t)c * a.P + a.slot) * M;
    float2 X[KPT];
#pragma unroll
    for (int q = 0; q < KPT; q++) {
        const int k = tid + q * NT;
        if (k < M) {
            X[q] = r2c_post_w<M>(Z, k, wn[q]);
            gst(row + k, X[q]);
        }
    }
    group_sync<NT>();

    // output samples 2n, 2n + 1 for n = tid + q * NT < M / 2
    float2 y[ZP];
#pragma unroll
    for (int f = 0; f < 2; f++) {
        if (f >= nf) continue;
        // the spectrum to invert, into s1; its C2R input into s0
#pragma unroll
        for (int q = 0; q < KPT; q++) {
            const int k = tid + q * NT;
            if (k < M) {
                float2 v = acc[f][q];
                mac2(v, X[q], h0[f][q], k == 0);
                sm.s1[k] = v;
            }
        }
        group_sync<NT>();
#pragma unroll
        for (int q = 0; q < KPT; q++) {
            const int k = tid + q * NT;
            if (k < M) sm.s0[k] = c2r_pre_w<M>(sm.s1, k, wn[q]);
        }
        group_sync<NT>();
        const float2 *z = fft_lds<M, true, NT>(sm.s0, sm.s1, sm.tw);
        // second half of the 2M output samples: z[M/2 .. M)
        if (f == 0) {
#pragma unroll
            for (int q = 0; q < ZP; q++) {
                const int n = tid + q * NT;
                if (n < M / 2) {
                    const float2 v = z[M / 2 + n];
                    y[q].x = fmaf(v.x, kScale, za[q].x);
                    y[q].y = fmaf(v.y, kScale, za[q].y);
                }
            }
        } else {
            const float step = (float)(1.0 / (double)M);
#pragma unroll
            for (int q = 0; q < ZP; q++) {
                const int n = tid + q * NT;
                if (n < M / 2) {
                    const float2 v = z[M / 2 + n];
                    const int j0 = 2 * n, j1 = 2 * n + 1;
                    const float y0 = fmaf(v.x, kScale, zb[q].x), y1 = fmaf(v.y, kScale, zb[q].y);
                    y[q].x = fmaf(y0, step * (float)j0, y[q].x * (step * (float)(M - j0)));
                    y[q].y = fmaf(y1, step * (float)j1, y[q].y * (step * (float)(M - j1)));
                }
            }
        }
        // the next filter's stores into s0 / s1 come after every lane's reads
        group_sync<NT>();
    }
    float *yo = a.yout + (size_t)c * a.out_stride;
#pragma unroll
    for (int q = 0; q < ZP; q++) {
        const int n = tid + q * NT;
        if (n < M / 2) {
            gst(yo + 2 * n, y[q].x);
            gst(yo + 2 * n + 1, y[q].y);
        }
    }
}

// Head workgroup b: channel b, or with WAVE the wave_heads(M) channels from
// wave_heads(M) * b, one per wave on its own HeadSmem.
template <int M, int HT, bool WAVE>
__device__ __forceinline__ void head_block(const TickArgs &a, int b, HeadSmem<M> *hs)
{
    constexpr int HPB = WAVE ? wave_heads(M) : 1;
    if constexpr (HPB > 1) {
        static_assert(HPB == kTickThreads / 64, "one head per wave");
        // the wave index as a uniform value: the channel's row addresses
        // stay in scalar registers
        const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), c = b * HPB + w;
        if (c >= a.nhead || !a.run[c]) return;  // wave-uniform: no workgroup barrier follows
        head_role<M, HT, 64>(a, c, hs[w]);
    } else {
        if (!a.run[b]) return;
        head_role<M, HT, kTickThreads>(a, b, *hs);
    }
}

// Rows of one channel for a partial sum over partitions [i0, i1): partition
// i multiplies FDL row (base - i) & mask with IR row i, over B bins starting
// at the rows' slice offsets.
struct MacView {
    const float4 *F;       // FDL row 0 of the channel, at the slice (float4 units)
    int fstride;           // float4 per FDL row
    int base, mask;
    const char *HA;        // IR row 0 of the channel, at the slice
    const char *HB;        // second filter or nullptr
    size_t hstride;        // bytes per IR row
    const float *SA;       // scales of the channel's rows (reduced precision)
    const float *SB;
    int i0, i1;
    bool dc;               // the slice starts with the packed (DC, Nyquist) element
    float4 *TA;            // this chunk's output row (B complex)
    float4 *TB;
};

template <int HT>
__device__ __forceinline__ float row_scale(const float *S, int i)
{
    if constexpr (HT == IR_F32) return 1.f;
    else return gld(S + i);
}

// 256 threads; a row of B complex is read as float4 (two complex) per lane,
// G rows in flight, U partitions unrolled per lane.
template <int M, int NF, int HT>
__device__ void mac_rows(const MacView &w, float4 *red)
{
    constexpr int H = M / 2;                     // float4 per row
    constexpr int V = H > 256 ? H / 256 : 1;     // float4 per lane per row
    constexpr int R = H / V;                     // lanes per row
    constexpr int G = kTickThreads / R;          // rows in flight
    constexpr int U = 8;                         // unroll (partitions per lane)
    const int tid = threadIdx.x;
    const int g = tid / R, r = tid % R;
    auto hrow = [&](const char *base, int i) {
        return reinterpret_cast<const void *>(base + (size_t)i * w.hstride);
    };

    float4 accA[V], accB[V];
#pragma unroll
    for (int v = 0; v < V; v++) {
        accA[v] = make_float4(0.f, 0.f, 0.f, 0.f);
        accB[v] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const bool dc0 = w.dc && (r == 0);
    const float s0 = dc0 ? 0.f : 1.f;

    int i = w.i0 + g;
    for (; i + (U - 1) * G < w.i1; i += U * G) {
        float4 xf[U][V], ha[U][V], hb[U][V];
        float sa[U], sb[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int ii = i + u * G;
            const int sl = (w.base - ii) & w.mask;
            sa[u] = row_scale<HT>(w.SA, ii);
            if (NF == 2) sb[u] = row_scale<HT>(w.SB, ii);
#pragma unroll
            for (int v = 0; v < V; v++) {
                xf[u][v] = ld_stream(w.F + (size_t)sl * w.fstride + v * R + r);
                ha[u][v] = IrRow<HT>::get2(hrow(w.HA, ii), v * R + r);
                if (NF == 2) hb[u][v] = IrRow<HT>::get2(hrow(w.HB, ii), v * R + r);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
#pragma unroll
            for (int v = 0; v < V; v++) {
                const bool dc = dc0 && v == 0;
                const float sv = v == 0 ? s0 : 1.f;
                mac4(accA[v], xf[u][v], HT == IR_F32 ? ha[u][v] : scale4(ha[u][v], sa[u]), sv, dc);
                if (NF == 2) mac4(accB[v], xf[u][v], HT == IR_F32 ? hb[u][v] : scale4(hb[u][v], sb[u]), sv, dc);
            }
        }
    }
    for (; i < w.i1; i += G) {
        const int sl = (w.base - i) & w.mask;
        const float sai = row_scale<HT>(w.SA, i);
        const float sbi = NF == 2 ? row_scale<HT>(w.SB, i) : 0.f;
#pragma unroll
        for (int v = 0; v < V; v++) {
            const bool dc = dc0 && v == 0;
            const float sv = v == 0 ? s0 : 1.f;
            const float4 x = ld_stream(w.F + (size_t)sl * w.fstride + v * R + r);
            const float4 h1 = IrRow<HT>::get2(hrow(w.HA, i), v * R + r);
            mac4(accA[v], x, HT == IR_F32 ? h1 : scale4(h1, sai), sv, dc);
            if (NF == 2) {
                const float4 h2 = IrRow<HT>::get2(hrow(w.HB, i), v * R + r);
                mac4(accB[v], x, HT == IR_F32 ? h2 : scale4(h2, sbi), sv, dc);
            }
        }
    }

    if constexpr (G == 1) {
#pragma unroll
        for (int v = 0; v < V; v++) {
            w.TA[v * R + r] = accA[v];
            if (NF == 2) w.TB[v * R + r] = accB[v];
        }
    } else {
        // red: [NF][G][H] float4
        static_assert(NF * G * H <= tail_red_float4<M>(), "tail reduction buffer");
#pragma unroll
        for (int v = 0; v < V; v++) {
            red[(0 * G + g) * H + v * R + r] = accA[v];
            if (NF == 2) red[(1 * G + g) * H + v * R + r] = accB[v];
        }
        __syncthreads();
        for (int e = tid; e < NF * H; e += kTickThreads) {
            const int f = e / H, q = e % H;
            float4 sum = red[(f * G + 0) * H + q];
#pragma unroll
            for (int gg = 1; gg < G; gg++) {
                const float4 t = red[(f * G + gg) * H + q];
                sum.x += t.x; sum.y += t.y; sum.z += t.z; sum.w += t.w;
            }
            (f == 0 ? w.TA : w.TB)[q] = sum;
        }
    }
}

// One lane group's share of a split pass: outputs [OFF, OFF + KH) over every
// partition of the chunk, the window and accumulators of KH outputs per lane;
// output OFF + k reads FDL row (tbase + OFF + k - i), so the loop is the
// pass's own with base and gap moved by OFF.  Sums go out directly.
template <int M, int NF, int HT, bool NT, int KH, int OFF>
__device__ void tailk_half(const TickArgs &a, int c, int j)
{
    constexpr int H = M / 2;                     // float4 per row
    constexpr int V = H > 256 ? H / 256 : 1;     // float4 per lane per row
    constexpr int R = H / V;                     // lanes per row
    constexpr int U = 4;                         // partitions loaded together (8 measured slower)
    static_assert(KH >= 1 && KH <= kPassMax, "pass length");
    constexpr size_t kRowBytes = (size_t)M * IrRow<HT>::kBytes;
    const int tid = threadIdx.x;
    const int r = tid % R;
    const float4 *F = reinterpret_cast<const float4 *>(a.fdl) + (size_t)c * a.P * H;
    const char *HA = reinterpret_cast<const char *>(a.tHa) + (size_t)c * a.P * kRowBytes;
    const char *HB = NF == 2 ? reinterpret_cast<const char *>(a.tHb) + (size_t)c * a.P * kRowBytes : nullptr;
    const float *SA = HT == IR_F32 ? nullptr : a.tSa + (size_t)c * a.P;
    const float *SB = (HT == IR_F32 || NF == 1) ? nullptr : a.tSb + (size_t)c * a.P;
    const int mask = a.P - 1, base = a.tbase + OFF, kn = a.kn - OFF, gap = a.kgap + OFF;
    const int i0 = 2 + j * a.pc, i1 = min(a.P, i0 + a.pc);
    const int p0 = i0, p1 = i1;
    bool wantA[KH], wantB[KH];
#pragma unroll
    for (int k = 0; k < KH; k++) {
        wantA[k] = k < kn && a.kTA[OFF + k] != nullptr;
        wantB[k] = NF == 2 && k < kn && a.kTB[OFF + k] != nullptr;
    }
    if (kn <= 0) return;  // a pass of at most OFF outputs: none of this group's
    const bool dc0 = (r == 0);
    const float s0 = dc0 ? 0.f : 1.f;
    auto frow = [&](int row, int v) { return ld_row<NT>(F + (size_t)(row & mask) * H + v * R + r); };
    auto hrow2 = [&](const char *row, int q) { return NT ? IrRow<HT>::get2(row, q) : IrRow<HT>::getc(row, q); };

    float4 accA[KH][V], accB[KH][V], xw[KH][V];
#pragma unroll
    for (int k = 0; k < KH; k++)
#pragma unroll
        for (int v = 0; v < V; v++) {
            accA[k][v] = make_float4(0.f, 0.f, 0.f, 0.f);
            accB[k][v] = make_float4(0.f, 0.f, 0.f, 0.f);
            // the window as it stands before partition p0: xw[k] = row (base + k + 1 - p0)
            // (output k + 1 reads it at p0), loaded only where some output uses it
            xw[k][v] = (k + 1 < KH && p0 < p1 && p0 >= k + 2 + gap) ? frow(base + k + 1 - p0, v)
                                                                     : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    for (int i = p0; i < p1; i += U) {
        float4 xn[U][V], ha[U][V], hb[U][V];
        float sa[U], sb[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int ii = i + u;
            sa[u] = 1.f;
            sb[u] = 1.f;
            if (ii < p1) {
                sa[u] = row_scale<HT>(SA, ii);
                if (NF == 2) sb[u] = row_scale<HT>(SB, ii);
            }
#pragma unroll
            for (int v = 0; v < V; v++) {
                if (ii < p1) {
                    xn[u][v] = frow(base - ii, v);
                    ha[u][v] = hrow2(HA + (size_t)ii * kRowBytes, v * R + r);
                    if (NF == 2) hb[u][v] = hrow2(HB + (size_t)ii * kRowBytes, v * R + r);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int ii = i + u;
            if (ii >= p1) break;
#pragma unroll
            for (int v = 0; v < V; v++) {
#pragma unroll
                for (int k = KH - 1; k > 0; k--) xw[k][v] = xw[k - 1][v];
                xw[0][v] = xn[u][v];
                const bool dc = dc0 && v == 0;
                const float sv = v == 0 ? s0 : 1.f;
                const float4 hA = HT == IR_F32 ? ha[u][v] : scale4(ha[u][v], sa[u]);
                float4 hB = hA;
                if (NF == 2) hB = HT == IR_F32 ? hb[u][v] : scale4(hb[u][v], sb[u]);
#pragma unroll
                for (int k = 0; k < KH; k++) {
                    if (ii >= k + 1 + gap) {
                        if (wantA[k]) mac4(accA[k][v], xw[k][v], hA, sv, dc);
                        if (NF == 2 && wantB[k]) mac4(accB[k][v], xw[k][v], hB, sv, dc);
                    }
                }
            }
        }
    }
    const size_t orow = ((size_t)c * a.nchunk + j) * H;
#pragma unroll
    for (int k = 0; k < KH; k++)
#pragma unroll
        for (int v = 0; v < V; v++) {
            if (wantA[k]) reinterpret_cast<float4 *>(a.kTA[OFF + k])[orow + v * R + r] = accA[k][v];
            if (NF == 2 && wantB[k]) reinterpret_cast<float4 *>(a.kTB[OFF + k])[orow + v * R + r] = accB[k][v];
        }
}

// Multi-block pass (kn >= 1 outputs): partitions [2 + j*pc, 2 + (j+1)*pc)
// of channel c for kn consecutive blocks.  Output k (0-based, block
// o + g + k, o the launch's first head block, g = kgap) takes partition
// i >= g + k + 1 with FDL row (tbase + k - i), the row partition i - k reads
// for output 0: each lane group walks its partitions in order and keeps the
// last kPassMax rows it loaded, so every IR row and FDL row is read once and
// used for every output.
template <int M, int NF, int HT, bool NT = true, int KM = kPassMax>
__device__ void tailk_role(const TickArgs &a, int c, int j, float4 *red)
{
    constexpr int H = M / 2;                     // float4 per row
    constexpr int V = H > 256 ? H / 256 : 1;     // float4 per lane per row
    constexpr int R = H / V;                     // lanes per row
    constexpr int G = kTickThreads / R;          // lane groups
    constexpr int U = 4;                         // partitions loaded together (8 measured slower)
    static_assert(KM >= 1 && KM <= kPassMax, "pass length");
    constexpr size_t kRowBytes = (size_t)M * IrRow<HT>::kBytes;
    const int tid = threadIdx.x;
    const int g = tid / R, r = tid % R;
    if constexpr (KM == kPassMax && G == 2 && R % 64 == 0) {
        // 16 outputs: each lane group takes 8 of them over every partition
        // (8 window rows and 8 accumulators per lane instead of 16, and no
        // reduction between the groups); the second group's FDL rows are the
        // first group's of 8 partitions before, its IR rows the same ones
        // (ordinary loads: a row must stay in the caches for the other group)
        if (__builtin_amdgcn_readfirstlane(g) == 0) tailk_half<M, NF, HT, false, kPassMax / 2, 0>(a, c, j);
        else tailk_half<M, NF, HT, false, kPassMax / 2, kPassMax / 2>(a, c, j);
        return;
    }
    const float4 *F = reinterpret_cast<const float4 *>(a.fdl) + (size_t)c * a.P * H;
    const char *HA = reinterpret_cast<const char *>(a.tHa) + (size_t)c * a.P * kRowBytes;
    const char *HB = NF == 2 ? reinterpret_cast<const char *>(a.tHb) + (size_t)c * a.P * kRowBytes : nullptr;
    const float *SA = HT == IR_F32 ? nullptr : a.tSa + (size_t)c * a.P;
    const float *SB = (HT == IR_F32 || NF == 1) ? nullptr : a.tSb + (size_t)c * a.P;
    const int mask = a.P - 1, base = a.tbase, kn = a.kn, gap = a.kgap;
    const int i0 = 2 + j * a.pc, i1 = min(a.P, i0 + a.pc);
    const int span = (i1 - i0 + G - 1) / G;
    const int p0 = min(i1, i0 + g * span), p1 = min(i1, p0 + span);
    bool wantA[KM], wantB[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        wantA[k] = k < kn && a.kTA[k] != nullptr;
        wantB[k] = NF == 2 && k < kn && a.kTB[k] != nullptr;
    }
    const bool dc0 = (r == 0);
    const float s0 = dc0 ? 0.f : 1.f;
    auto frow = [&](int row, int v) { return ld_row<NT>(F + (size_t)(row & mask) * H + v * R + r); };
    auto hrow2 = [&](const char *row, int q) { return NT ? IrRow<HT>::get2(row, q) : IrRow<HT>::getc(row, q); };

    float4 accA[KM][V], accB[KM][V], xw[KM][V];
#pragma unroll
    for (int k = 0; k < KM; k++)
#pragma unroll
        for (int v = 0; v < V; v++) {
            accA[k][v] = make_float4(0.f, 0.f, 0.f, 0.f);
            accB[k][v] = make_float4(0.f, 0.f, 0.f, 0.f);
            // the window as it stands before partition p0: xw[k] = row (base + k + 1 - p0)
            // (output k + 1 reads it at p0), loaded only where some output uses it
            xw[k][v] = (k + 1 < KM && p0 < p1 && p0 >= k + 2 + gap) ? frow(base + k + 1 - p0, v)
                                                                     : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    for (int i = p0; i < p1; i += U) {
        float4 xn[U][V], ha[U][V], hb[U][V];
        float sa[U], sb[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int ii = i + u;
            sa[u] = 1.f;
            sb[u] = 1.f;
            if (ii < p1) {
                sa[u] = row_scale<HT>(SA, ii);
                if (NF == 2) sb[u] = row_scale<HT>(SB, ii);
            }
#pragma unroll
            for (int v = 0; v < V; v++) {
                if (ii < p1) {
                    xn[u][v] = frow(base - ii, v);
                    ha[u][v] = hrow2(HA + (size_t)ii * kRowBytes, v * R + r);
                    if (NF == 2) hb[u][v] = hrow2(HB + (size_t)ii * kRowBytes, v * R + r);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int ii = i + u;
            if (ii >= p1) break;
#pragma unroll
            for (int v = 0; v < V; v++) {
#pragma unroll
                for (int k = KM - 1; k > 0; k--) xw[k][v] = xw[k - 1][v];
                xw[0][v] = xn[u][v];
                const bool dc = dc0 && v == 0;
                const float sv = v == 0 ? s0 : 1.f;
                const float4 hA = HT == IR_F32 ? ha[u][v] : scale4(ha[u][v], sa[u]);
                float4 hB = hA;
                if (NF == 2) hB = HT == IR_F32 ? hb[u][v] : scale4(hb[u][v], sb[u]);
#pragma unroll
                for (int k = 0; k < KM; k++) {
                    if (ii >= k + 1 + gap) {
                        if (wantA[k]) mac4(accA[k][v], xw[k][v], hA, sv, dc);
                        if (NF == 2 && wantB[k]) mac4(accB[k][v], xw[k][v], hB, sv, dc);
                    }
                }
            }
        }
    }
    // the lane groups' sums reduced through red one output at a time:
    // [NF][G][H] float4 per round
    const size_t orow = ((size_t)c * a.nchunk + j) * H;
    if constexpr (G == 1) {
#pragma unroll
        for (int k = 0; k < KM; k++)
#pragma unroll
            for (int v = 0; v < V; v++) {
                if (wantA[k]) reinterpret_cast<float4 *>(a.kTA[k])[orow + v * R + r] = accA[k][v];
                if (NF == 2 && wantB[k]) reinterpret_cast<float4 *>(a.kTB[k])[orow + v * R + r] = accB[k][v];
            }
    } else {
        static_assert(NF * G * H <= tail_red_float4<M>(), "pass reduction buffer");
        // unrolled over the outputs so that every pointer is picked
        // statically (an output pointer indexed at run time would put the
        // whole argument struct in private memory); the skip is uniform
#pragma unroll
        for (int k = 0; k < KM; k++) {
            if (!wantA[k] && !(NF == 2 && wantB[k])) continue;
#pragma unroll
            for (int v = 0; v < V; v++) {
                red[(0 * G + g) * H + v * R + r] = accA[k][v];
                if (NF == 2) red[(1 * G + g) * H + v * R + r] = accB[k][v];
            }
            __syncthreads();
            // one loop per filter: a pointer chosen between the two at run
            // time would also put the argument copy in private memory
            auto reduce_to = [&](int f, float2 *out) {
                for (int q = tid; q < H; q += kTickThreads) {
                    float4 sum = red[(f * G + 0) * H + q];
#pragma unroll
                    for (int gg = 1; gg < G; gg++) {
                        const float4 t = red[(f * G + gg) * H + q];
                        sum.x += t.x; sum.y += t.y; sum.z += t.z; sum.w += t.w;
                    }
                    reinterpret_cast<float4 *>(out)[orow + q] = sum;
                }
            };
            if (wantA[k]) reduce_to(0, a.kTA[k]);
            if (NF == 2 && wantB[k]) reduce_to(1, a.kTB[k]);
            __syncthreads();
        }
    }
}

// Uniform tail: partitions [2 + j*pc, 2 + (j+1)*pc) of channel c for the next block.
template <int M, int NF, int HT>
__device__ void tail_role(const TickArgs &a, int c, int j, float4 *red)
{
    constexpr int H = M / 2;
    constexpr size_t kRowBytes = (size_t)M * IrRow<HT>::kBytes;
    MacView w;
    w.F = reinterpret_cast<const float4 *>(a.fdl) + (size_t)c * a.P * H;
    w.fstride = H;
    w.base = a.tbase;
    w.mask = a.P - 1;
    w.HA = reinterpret_cast<const char *>(a.tHa) + (size_t)c * a.P * kRowBytes;
    w.HB = NF == 2 ? reinterpret_cast<const char *>(a.tHb) + (size_t)c * a.P * kRowBytes : nullptr;
    w.hstride = kRowBytes;
    w.SA = HT == IR_F32 ? nullptr : a.tSa + (size_t)c * a.P;
    w.SB = (HT == IR_F32 || NF == 1) ? nullptr : a.tSb + (size_t)c * a.P;
    w.i0 = 2 + j * a.pc;
    w.i1 = min(a.P, w.i0 + a.pc);
    w.dc = true;
    w.TA = reinterpret_cast<float4 *>(a.tTa) + ((size_t)c * a.nchunk + j) * H;
    w.TB = NF == 2 ? reinterpret_cast<float4 *>(a.tTb) + ((size_t)c * a.nchunk + j) * H : nullptr;
    mac_rows<M, NF, HT>(w, red);
}

// Frame multiply-accumulate of a non-uniform object: bin slice a.bslice of
// frame partitions [2 + k*pcB, 2 + (k+1)*pcB) for channel c.
template <int M, int NF, int HT>
__device__ void frame_mac_role(const TickArgs &a, int c, int k, float4 *red)
{
    const size_t rowB = (size_t)a.N * IrRow<HT>::kBytes;
    const size_t off = (size_t)a.bslice * M;  // complex elements
    MacView w;
    w.F = reinterpret_cast<const float4 *>(a.fb + (size_t)c * a.RB * a.N + off);
    w.fstride = a.N / 2;
    w.base = a.fbase;
    w.mask = a.RB - 1;
    w.HA = reinterpret_cast<const char *>(a.bHa) + (size_t)c * a.PB * rowB + off * IrRow<HT>::kBytes;
    w.HB = NF == 2 ? reinterpret_cast<const char *>(a.bHb) + (size_t)c * a.PB * rowB + off * IrRow<HT>::kBytes
                   : nullptr;
    w.hstride = rowB;
    w.SA = HT == IR_F32 ? nullptr : a.bSa + (size_t)c * a.PB;
    w.SB = (HT == IR_F32 || NF == 1) ? nullptr : a.bSb + (size_t)c * a.PB;
    w.i0 = 2 + k * a.pcB;
    w.i1 = min(a.PB, w.i0 + a.pcB);
    w.dc = a.bslice == 0;
    w.TA = reinterpret_cast<float4 *>(a.bTa + ((size_t)c * a.nchB + k) * a.N + off);
    w.TB = NF == 2 ? reinterpret_cast<float4 *>(a.bTb + ((size_t)c * a.nchB + k) * a.N + off) : nullptr;
    mac_rows<M, NF, HT>(w, red);
}

// frame multiply-accumulate blocks (after head and tail blocks)
template <int M, int HT>
__device__ __forceinline__ void frame_mac_block(const TickArgs &a, int t, float4 *smem)
{
    if (t >= a.C * a.nchB) return;
    const int c = t / a.nchB, k = t % a.nchB;
    if (!a.run[c]) return;
    if (a.bHb) frame_mac_role<M, 2, HT>(a, c, k, smem);
    else frame_mac_role<M, 1, HT>(a, c, k, smem);
}

// TK: 0 = a launch without tail blocks (head and frame roles only),
// 1 = the next block's tail, 2 = a multi-block pass.  Separate
// instantiations keep head-only launches light and let profiles tell the
// launch kinds apart.
template <int M, int NF, int HT, int TK, int KM = kPassMax>
__global__ __launch_bounds__(kTickThreads) void k_tick(TickArgs a)
{
    // tail or pass reduction buffer (float4): [NF][G][H]
    constexpr int kTailRed = tail_red_float4<M>();
    constexpr int kHeadBytes = sizeof(HeadSmem<M>);
    constexpr int kBytes = kHeadBytes > kTailRed * 16 ? kHeadBytes : kTailRed * 16;
    __shared__ float4 smem[(kBytes + 15) / 16];
    const int b = blockIdx.x;
    if (b < a.nhead) {
        head_block<M, HT, false>(a, b, reinterpret_cast<HeadSmem<M> *>(smem));
    } else if (TK != 0 && b < a.nhead + a.C * a.nchunk) {
        const int t = b - a.nhead;
        const int c = t / a.nchunk, j = t % a.nchunk;
        if (!a.run[c]) return;
        if constexpr (TK == 2) tailk_role<M, NF, HT, true, KM>(a, c, j, smem);
        else tail_role<M, NF, HT>(a, c, j, smem);
    } else {
        frame_mac_block<M, HT>(a, b - a.nhead - a.C * a.nchunk, smem);
    }
}

// Roles of one tick for block index b of one object's launch.  Role sets of
// a tick launch (each compiles only its roles, so a launch carries the
// registers of what it runs): kRolesGeneral: heads, single tails and frame
// slices (objects without multi-block passes); kRolesLean: heads and
// single-filter passes of at most KM outputs; kRolesXfade: heads and
// two-filter passes (a switch in flight).
template <int M, int HT, int ROLES, int KM = kPassMax>
__device__ __forceinline__ void tick_body(const TickArgs &a, int b, float4 *smem)
{
    if (b < a.nhead) {
        head_block<M, HT, false>(a, b, reinterpret_cast<HeadSmem<M> *>(smem));
    } else {
        const int t = b - a.nhead;
        if (t >= a.C * a.nchunk) {
            if constexpr (ROLES == kRolesGeneral) frame_mac_block<M, HT>(a, t - a.C * a.nchunk, smem);
            return;
        }
        const int c = t / a.nchunk, j = t % a.nchunk;
        if (!a.run[c]) return;
        if constexpr (ROLES == kRolesLean) {
            tailk_role<M, 1, HT, true, KM>(a, c, j, smem);
        } else if constexpr (ROLES == kRolesXfade) {
            tailk_role<M, 2, HT, true, KM>(a, c, j, smem);
        } else {
            if (a.tHb) tail_role<M, 2, HT>(a, c, j, smem);
            else tail_role<M, 1, HT>(a, c, j, smem);
        }
    }
}

// Graph-replayed variant: the arguments of launch k come from a descriptor
// array in device memory (written by one copy per batch), so a captured
// graph of k nodes can be replayed with new arguments; the tail's filter
// count is read from the descriptor.  Captured objects run a tail every
// block (no multi-block passes).
template <int M, int HT>
__global__ __launch_bounds__(kTickThreads) void k_tick_desc(const TickArgs *__restrict__ descs, int k)
{
    constexpr int kTailRed = tail_red_float4<M>();  // float4
    constexpr int kHeadBytes = sizeof(HeadSmem<M>);
    constexpr int kBytes = kHeadBytes > kTailRed * 16 ? kHeadBytes : kTailRed * 16;
    __shared__ float4 smem[(kBytes + 15) / 16];
    tick_body<M, HT, kRolesGeneral>(descs[k], blockIdx.x, smem);
}

// One launch for the objects of one block size M and role set (a group
// call launches one per block size, role set and pass length present).  A
// workgroup finds its object by binary search over the block offsets.
template <int HT, int M, int ROLES, int KM>
__global__ __launch_bounds__(kTickThreads) void k_group_tick(const TickArgs *__restrict__ descs,
                                                             const int *__restrict__ starts, int nobj)
{
    // tail or pass reduction buffer (float4): [NF][G][H]
    constexpr int kTailRed = tail_red_float4<M>();
    constexpr int kHeadBytes = sizeof(HeadSmem<M>);
    constexpr int kBytes = kHeadBytes > kTailRed * 16 ? kHeadBytes : kTailRed * 16;
    __shared__ float4 smem[(kBytes + 15) / 16];
    const int b = blockIdx.x;
    int lo = 0, hi = nobj - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    // a copy: its fields load once, before any store, into scalar registers
    // (read in place after the roles' stores, they would be vector loads)
    const TickArgs a = descs[lo];
    if (a.B != M) return;  // the host launches each object with its own block size
    tick_body<M, HT, ROLES, KM>(a, b - starts[lo], smem);
}

// Heads-only group launch: the objects of block size M without a pass, tail
// or frame slice this tick, wave_heads(M) heads per workgroup.  Without the
// pass roles' registers and with four heads per workgroup the heads' FFT
// chains overlap four times as many memory round trips per compute unit;
// at least 4 waves per SIMD (at most 128 VGPRs) keep 16 heads per compute
// unit in flight.
template <int HT, int M>
__global__ __launch_bounds__(kTickThreads, 4) void k_group_heads(const TickArgs *__restrict__ descs,
                                                                 const int *__restrict__ starts, int nobj)
{
    __shared__ HeadSmem<M> sm[wave_heads(M)];
    const int b = blockIdx.x;
    int lo = 0, hi = nobj - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const TickArgs a = descs[lo];
    if (a.B != M) return;
    head_block<M, HT, true>(a, b - starts[lo], sm);
}

// Work-queue form of a group launch: a grid of the device's resident
// workgroups takes the launch's workgroup items [0, total) in order from
// *next (zero before the launch) and runs each one as k_group_tick or, for
// the heads kind, k_group_heads would run that workgroup (the same roles,
// so the same bits).  Every item is in the descriptors before the launch:
// a workgroup leaves once the counter has passed total, waiting on nothing.
template <int HT, int M, int ROLES, int KM>
__global__ __launch_bounds__(kTickThreads) void k_group_queue(const TickArgs *__restrict__ descs,
                                                              const int *__restrict__ starts, int nobj, int total,
                                                              int *__restrict__ next)
{
    constexpr int kTailRed = tail_red_float4<M>();
    constexpr int kHeadBytes = (ROLES == kRolesHeads ? wave_heads(M) : 1) * (int)sizeof(HeadSmem<M>);
    constexpr int kBytes = kHeadBytes > kTailRed * 16 ? kHeadBytes : kTailRed * 16;
    __shared__ float4 smem[(kBytes + 15) / 16];
    __shared__ int item;
    for (;;) {
        if (threadIdx.x == 0) item = atomicAdd(next, 1);
        __syncthreads();
        const int b = item;
        __syncthreads();  // every thread has read item before the next one is taken
        if (b >= total) break;
        int lo = 0, hi = nobj - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (starts[mid] <= b) lo = mid;
            else hi = mid - 1;
        }
        const TickArgs a = descs[lo];
        if (a.B == M) {
            if constexpr (ROLES == kRolesHeads)
                head_block<M, HT, true>(a, b - starts[lo], reinterpret_cast<HeadSmem<M> *>(smem));
            else
                tick_body<M, HT, ROLES, KM>(a, b - starts[lo], smem);
        }
        __syncthreads();  // the item's LDS use is over before the next item
    }
}

// ---------------------------------------------------------------------------
// Batched head-only blocks (HeadBatchArgs): one workgroup per channel runs
// the heads of n consecutive blocks.  Their forward and inverse FFTs are
// independent and share the stages (and barriers) of one; only block j's
// multiply-accumulate reads the spectra of blocks j-1, ..., 0 of the batch
// (from LDS) and of the blocks before the batch (from the FDL).  Partitions
// are added in the order of head_role (partials, d..1, then 0), so a batch
// gives the same bits as n single-block launches.
// ---------------------------------------------------------------------------
constexpr int kHeadsChunkBatch = 8;  // partial-sum rows per block loaded together by k_heads

template <int M>
struct HeadsSmem {
    float2 a[kHeadBatch][M], b[kHeadBatch][M], x[kHeadBatch][M], tw[M], twn[M];
};

template <int M, int HT, int NEAR = kNearMax>
__device__ void heads_role(const HeadBatchArgs &h, int c, HeadsSmem<M> &sm)
{
    constexpr int NT = kTickThreads;
    constexpr int KPT = M >= NT ? M / NT : 1;  // bins per thread
    constexpr int LM = ilog2(M);
    constexpr float kScale = 1.0f / (2.0f * (float)M);
    const int tid = threadIdx.x, n = h.n;
    if (!h.run[c]) return;
    // the batch is the launch's longest chain: its waves issue before the
    // pass's on a shared compute unit
    __builtin_amdgcn_s_setprio(3);

    // frames [x_{t-1}, x_t] of every entry as complex pairs (entries past
    // the batch repeat its last block: every load is in range and they all
    // go out together), and the twiddles
#pragma unroll
    for (int j = 0; j < kHeadBatch; j++) {
        const float *prev = j == 0 ? h.xprev + (size_t)c * M : h.xin[j - 1] + (size_t)c * h.in_stride;
        const float *cur = h.xin[j] + (size_t)c * h.in_stride;
#pragma unroll
        for (int m = tid; m < M; m += NT) {
            const float *src = m < M / 2 ? prev + 2 * m : cur + 2 * (m - M / 2);
            sm.a[j][m] = make_float2(src[0], src[1]);
        }
    }
    for (int m = tid; m < M; m += NT) {
        sm.tw[m] = h.twM[m];
        sm.twn[m] = h.twN[m];
    }
    // Every global load of the batch is issued before any is used (one
    // memory round trip).  The blocks of a batch share one filter (no
    // crossfade inside a batch) and the FDL rows before it, so a thread
    // holds, for its bins, IR partitions 0..NEAR and the rows of blocks
    // w-1 .. w-NEAR once (clamped indices, used where in range); block
    // j adds partitions d_j..j+1 from those rows before its FFT and
    // partitions j..0 from the batch's spectra after it.  Entries j >= n
    // repeat block n-1 (valid addresses, results unused).
    constexpr int KP = kHeadBatch;
    float2 acc[KP][KPT], hr[NEAR + 1][KPT], fr[NEAR + 1][KPT];
    const int dlim = min(NEAR, h.P - 1);
#pragma unroll
    for (int q = 0; q < KPT; q++) {
        const int k = min(tid + q * NT, M - 1);
#pragma unroll
        for (int i = 0; i <= NEAR; i++) {
            const int ii = min(i, dlim);
            const float si = ir_scale<HT>(h.S[0], c, h.P, ii);
            const float2 hv = IrRow<HT>::get1(ir_row<HT>(h.H[0], c, h.P, ii, M), k);
            hr[i][q] = make_float2(hv.x * si, hv.y * si);
            // row of block w - i (i >= 1), at ring slot slot_0 - i
            fr[i][q] = h.fdl[((size_t)c * h.P + ((h.slot[0] - max(ii, 1)) & (h.P - 1))) * M + k];
        }
    }
    int tmax = 0;
#pragma unroll
    for (int j = 0; j < KP; j++) tmax = max(tmax, h.tchunk[j]);
#pragma unroll
    for (int q = 0; q < KPT; q++)
#pragma unroll
        for (int j = 0; j < KP; j++) acc[j][q] = make_float2(0.f, 0.f);
    for (int u0 = 0; u0 < tmax; u0 += kHeadsChunkBatch) {
#pragma unroll
        for (int q = 0; q < KPT; q++) {
            const int k = tid + q * NT;
            if (k >= M) continue;
            float2 t[KP][kHeadsChunkBatch];
#pragma unroll
            for (int j = 0; j < KP; j++) {
                const float2 *T = h.T[j] + ((size_t)c * h.tchunk[j] + u0) * M + k;
#pragma unroll
                for (int u = 0; u < kHeadsChunkBatch; u++)
                    t[j][u] = u0 + u < h.tchunk[j] ? T[(size_t)u * M] : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int j = 0; j < KP; j++)
#pragma unroll
                for (int u = 0; u < kHeadsChunkBatch; u++) {
                    acc[j][q].x += t[j][u].x;
                    acc[j][q].y += t[j][u].y;
                }
        }
    }
#pragma unroll
    for (int q = 0; q < KPT; q++) {
        const int k = tid + q * NT;
        if (k >= M) continue;
#pragma unroll
        for (int j = 0; j < KP; j++) {
            const int d = min(h.d[j], dlim);
            // partitions d..j+1: rows of blocks w-1 .. (rows before the batch), the larger first
#pragma unroll
            for (int i = NEAR; i > j; i--)
                if (i <= d) mac2(acc[j][q], fr[i - j][q], hr[i][q], k == 0);
        }
    }
    __syncthreads();
    {
        // the batch's last block becomes the previous one of the next launch
        const float *xl = h.xin[n - 1] + (size_t)c * h.in_stride;
        float *xpw = h.xprev + (size_t)c * M;
        for (int m = tid; m < M; m += NT) xpw[m] = xl[m];
    }

    const float2 *Z = fft_lds_n<M, false>(&sm.a[0][0], &sm.b[0][0], n, sm.tw);
    for (int e = tid; e < (n << LM); e += NT) {
        const int j = e >> LM, k = e & (M - 1);
        const float2 X = r2c_post<M>(Z + j * M, k, sm.twn);
        sm.x[j][k] = X;
        h.fdl[((size_t)c * h.P + h.slot[j]) * M + k] = X;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kHeadBatch; j++) {
        if (j >= n) continue;
        const int d = min(h.d[j], h.P - 1);
#pragma unroll
        for (int q = 0; q < KPT; q++) {
            const int k = tid + q * NT;
            if (k < M) {
                float2 v = acc[j][q];
#pragma unroll
                for (int i = KP - 1; i >= 0; i--)
                    if (i <= j && i <= d) mac2(v, sm.x[j - i][k], hr[i][q], k == 0);
                sm.a[j][k] = v;
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < (n << LM); e += NT) {
        const int j = e >> LM, k = e & (M - 1);
        sm.b[j][k] = c2r_pre<M>(&sm.a[j][0], k, sm.twn);
    }
    __syncthreads();
    const float2 *z = fft_lds_n<M, true>(&sm.b[0][0], &sm.a[0][0], n, sm.tw);
    // second half of each block's 2M output samples: z[M/2 .. M)
    for (int e = tid; e < (n << (LM - 1)); e += NT) {
        const int j = e >> (LM - 1), m = e & (M / 2 - 1);
        const float2 v = z[j * M + M / 2 + m];
        float *yo = h.yout[j] + (size_t)c * h.out_stride;
        yo[2 * m] = v.x * kScale;
        yo[2 * m + 1] = v.y * kScale;
    }
}

// NEAR: the most partitions a head of the batch adds (kHeadBatch + the
// object's pass length - 1)
template <int M, int HT, int NEAR>
__global__ __launch_bounds__(kTickThreads) void k_heads(HeadBatchArgs h)
{
    __shared__ HeadsSmem<M> sm;
    heads_role<M, HT, NEAR>(h, blockIdx.x, sm);
}

// A head batch and the pass for the blocks after it in one launch: blocks
// [0, C) run the heads (started first: their chain of FFTs is the longer
// one), the rest the pass, which reads FDL rows before the batch only.
template <int M, int NF, int HT, int KM>
__global__ __launch_bounds__(kTickThreads) void k_window(HeadBatchArgs h, TickArgs a)
{
    constexpr int kTailRed = tail_red_float4<M>();  // float4
    constexpr int kBytes = sizeof(HeadsSmem<M>) > kTailRed * 16 ? sizeof(HeadsSmem<M>) : kTailRed * 16;
    __shared__ float4 smem[(kBytes + 15) / 16];
    const int b = blockIdx.x;
    if (b < h.C) {
        heads_role<M, HT, kHeadBatch + KM - 1>(h, b, *reinterpret_cast<HeadsSmem<M> *>(smem));
    } else {
        const int t = b - h.C;
        const int c = t / a.nchunk, j = t % a.nchunk;
        if (!a.run[c]) return;
        if (a.cached) tailk_role<M, NF, HT, false, KM>(a, c, j, smem);
        else tailk_role<M, NF, HT, true, KM>(a, c, j, smem);
    }
}

// ---------------------------------------------------------------------------
// Frame kernel of non-uniform objects (see FrameArgs), 256 threads.
// ---------------------------------------------------------------------------
// threads of a frame workgroup: one radix-4 butterfly per thread and stage
template <int N>
constexpr int frame_threads() { return N / 4 < 256 ? 256 : (N / 4 > 1024 ? 1024 : N / 4); }
constexpr int kGroupFrameThreads = 1024;

template <int N, int NT, int HT>
__device__ void frame_role(const FrameArgs &a, int c, float2 *s0, float2 *s1)
{
    constexpr int KPT = N >= NT ? N / NT : 1;
    constexpr float kScale = 1.0f / (2.0f * (float)N);
    const int tid = threadIdx.x;
    const int nf = a.Hb ? 2 : 1;
    const int lB = __builtin_ctz((unsigned)a.B);
    const int rmask = 2 * a.M - 1;

    // [x_{g-1}, x_g]: the 2M ring blocks after the last one written, oldest first
    for (int n = tid; n < N; n += NT) {
        const int t = 2 * n;
        const int kb = t >> lB, w = t & (a.B - 1);
        const int sl = (a.xlast + 1 + kb) & rmask;
        const float2 v = *reinterpret_cast<const float2 *>(a.xring + ((size_t)c * 2 * a.M + sl) * a.B + w);
        s0[n] = v;
    }
    __syncthreads();
    const float2 *Z = fft_lds<N, false>(s0, s1, a.twN);
    float2 *zbuf = const_cast<float2 *>(Z);
    float2 *fr = (Z == s0) ? s1 : s0;

    float2 X[KPT], Xp[KPT];
    float2 *row = a.fb + ((size_t)c * a.RB + a.frow) * N;
    const float2 *prow = a.fb + ((size_t)c * a.RB + ((a.frow - 1) & (a.RB - 1))) * N;
#pragma unroll
    for (int q = 0; q < KPT; q++) {
        const int k = tid + q * NT;
        X[q] = make_float2(0.f, 0.f);
        Xp[q] = make_float2(0.f, 0.f);
        if (k < N) {
            X[q] = r2c_post<N>(Z, k, a.tw2N);
            row[k] = X[q];
            if (a.PB > 1) Xp[q] = prow[k];
        }
    }
    constexpr size_t kRowBytes = (size_t)N * IrRow<HT>::kBytes;
#pragma unroll
    for (int f = 0; f < 2; f++) {
        if (f >= nf) continue;
        const char *Hc = reinterpret_cast<const char *>(f == 0 ? a.Ha : a.Hb) + (size_t)c * a.PB * kRowBytes;
        const float *Sc = HT == IR_F32 ? nullptr : (f == 0 ? a.Sa : a.Sb) + (size_t)c * a.PB;
        const float sc0 = row_scale<HT>(Sc, 0);
        const float sc1 = a.PB > 1 ? row_scale<HT>(Sc, 1) : 0.f;
        const float2 *T = (f == 0 ? a.Ta : a.Tb) + (size_t)c * a.nch * N;
#pragma unroll
        for (int q = 0; q < KPT; q++) {
            const int k = tid + q * NT;
            if (k < N) {
                float2 acc = make_float2(0.f, 0.f);
                for (int j = 0; j < a.nch; j++) {
                    const float2 t = T[(size_t)j * N + k];
                    acc.x += t.x;
                    acc.y += t.y;
                }
                if (a.PB > 1) {
                    const float2 h1 = IrRow<HT>::get1(Hc + kRowBytes, k);
                    mac2(acc, Xp[q], make_float2(h1.x * sc1, h1.y * sc1), k == 0);
                }
                const float2 h0 = IrRow<HT>::get1(Hc, k);
                mac2(acc, X[q], make_float2(h0.x * sc0, h0.y * sc0), k == 0);
                fr[k] = acc;
            }
        }
        __syncthreads();
        for (int k = tid; k < N; k += NT) zbuf[k] = c2r_pre<N>(fr, k, a.tw2N);
        __syncthreads();
        const float2 *z = fft_lds<N, true>(zbuf, fr, a.twN);
        float *out = (f == 0 ? a.za : a.zb) + (size_t)c * N;
        for (int n = tid; n < N / 2; n += NT) {
            const float2 v = z[N / 2 + n];
            out[2 * n] = v.x * kScale;
            out[2 * n + 1] = v.y * kScale;
        }
        __syncthreads();
    }
}

template <int N, int HT>
__global__ __launch_bounds__(frame_threads<N>()) void k_frame(FrameArgs a)
{
    __shared__ float2 s0[N], s1[N];
    const int c = blockIdx.x;
    if (!a.run[c]) return;
    frame_role<N, frame_threads<N>(), HT>(a, c, s0, s1);
}

template <int HT>
__global__ __launch_bounds__(kGroupFrameThreads) void k_group_frame(const FrameArgs *__restrict__ descs,
                                                                    const int *__restrict__ starts, int nobj)
{
    __shared__ float2 s0[kMaxFrame], s1[kMaxFrame];
    const int b = blockIdx.x;
    int lo = 0, hi = nobj - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const FrameArgs a = descs[lo];
    const int c = b - starts[lo];
    if (c >= a.C || !a.run[c]) return;
    constexpr int NT = kGroupFrameThreads;
    switch (a.N) {
    case 128: frame_role<128, NT, HT>(a, c, s0, s1); break;
    case 256: frame_role<256, NT, HT>(a, c, s0, s1); break;
    case 512: frame_role<512, NT, HT>(a, c, s0, s1); break;
    case 1024: frame_role<1024, NT, HT>(a, c, s0, s1); break;
    case 2048: frame_role<2048, NT, HT>(a, c, s0, s1); break;
    case 4096: frame_role<4096, NT, HT>(a, c, s0, s1); break;
    default: break;
    }
}

// ---------------------------------------------------------------------------
// IR partition transform: one workgroup per (partition i, listed channel).
// Partition i of channel ch becomes the packed spectrum of [h[iB..iB+B), 0_B].
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned short f2bf_rne(float f)
{
    const unsigned int u = __float_as_uint(f);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// Partition i of list entry y (channel a.chan[y], or y) of one IR.
template <int M, int HT>
__device__ __forceinline__ void ir_transform_part(const IrArgs &a, int i, int y, float2 *s0, float2 *s1,
                                                  float *smax)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    const int ch = a.chan ? gld(a.chan + y) : y;
    const float *h = a.ir + (size_t)ch * a.stride;
    const long long base = (long long)i * M;
    for (int n = tid; n < M; n += nt) {
        float2 v = make_float2(0.f, 0.f);
        if (n < M / 2) {
            const long long k = base + 2 * n;
            v.x = k < a.n ? gld(h + k) : 0.f;
            v.y = k + 1 < a.n ? gld(h + k + 1) : 0.f;
        }
        s0[n] = v;
    }
    __syncthreads();
    const float2 *Z = fft_lds<M, false>(s0, s1, a.twM);
    if constexpr (HT == IR_F32) {
        float2 *row = reinterpret_cast<float2 *>(a.H) + ((size_t)ch * a.P + i) * M;
        for (int k = tid; k < M; k += nt) gst(row + k, r2c_post<M>(Z, k, a.twN));
    } else {
        // spectrum into the free buffer, then the row's max |re|,|im|
        float2 *X = (Z == s0) ? s1 : s0;
        float m = 0.f;
        for (int k = tid; k < M; k += nt) {
            const float2 v = r2c_post<M>(Z, k, a.twN);
            X[k] = v;
            m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
        }
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if ((tid & 63) == 0) smax[tid >> 6] = m;
        __syncthreads();
        float sc = 0.f;
        for (int w = 0; w < (nt + 63) / 64; w++) sc = fmaxf(sc, smax[w]);
        if (sc == 0.f) sc = 1.f;
        const float inv = 1.0f / sc;
        unsigned int *row = reinterpret_cast<unsigned int *>(a.H) + ((size_t)ch * a.P + i) * M;
        for (int k = tid; k < M; k += nt) {
            const float2 v = X[k];
            unsigned int lo, hi;
            if constexpr (HT == IR_F16) {
                lo = __half_as_ushort(__float2half(v.x * inv));
                hi = __half_as_ushort(__float2half(v.y * inv));
            } else {
                lo = f2bf_rne(v.x * inv);
                hi = f2bf_rne(v.y * inv);
            }
            gst(row + k, lo | (hi << 16));
        }
        if (tid == 0) gst(a.S + (size_t)ch * a.P + i, sc);
    }
}

template <int M, int HT>
__global__ __launch_bounds__(M / 4 > 64 ? M / 4 : 64) void k_ir_transform(IrArgs a)
{
    __shared__ float2 s0[M], s1[M];
    __shared__ float smax[32];
    ir_transform_part<M, HT>(a, blockIdx.x, blockIdx.y, s0, s1, smax);
}

// Entry j of a descriptor list whose block offsets start at starts[j]
__device__ __forceinline__ int group_entry(const int *__restrict__ starts, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Several IRs' transforms (one entry each: every partition of its listed
// channels, P blocks per channel) in one launch.
template <int M, int HT>
__global__ __launch_bounds__(M / 4 > 64 ? M / 4 : 64) void k_ir_transform_group(const IrArgs *__restrict__ descs,
                                                                                const int *__restrict__ starts, int n)
{
    __shared__ float2 s0[M], s1[M];
    __shared__ float smax[32];
    const int b = blockIdx.x, j = group_entry(starts, n, b);
    const IrArgs a = descs[j];
    const int local = b - starts[j];
    ir_transform_part<M, HT>(a, local % a.P, local / a.P, s0, s1, smax);
}

// FP32 IR spectra of NR partitions per workgroup (256 threads): the rows'
// FFTs share the Stockham stages, so a transform of many short partitions
// runs as a quarter (or half) as many workgroups.  The twiddles are staged
// in LDS with the samples (one memory round trip before the stages), and
// aligned rows load four samples per lane.  Same arithmetic per row as
// k_ir_transform.
template <int M, int NR>
__device__ __forceinline__ void ir_transform_rows_part(const IrArgs &a, int xb, int y, float2 *s0, float2 *s1,
                                                       float2 *tw, float2 *twn)
{
    const int i0 = xb * NR, tid = threadIdx.x, nt = blockDim.x;
    const int ch = a.chan ? gld(a.chan + y) : y;
    const int nr = min(NR, a.P - i0);
    const float *h = a.ir + (size_t)ch * a.stride;
    const bool al = (reinterpret_cast<uintptr_t>(h) & 15) == 0;
    // row r holds [h[(i0+r)M .. +M), 0_M] as M complex pairs: lane e loads
    // samples 4q..4q+3 of row r and zeroes the matching pairs of the upper half
    for (int e = tid; e < NR * (M / 4); e += nt) {
        const int r = e / (M / 4), q = e - r * (M / 4);
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (r < nr) {
            const long long k = (long long)(i0 + r) * M + 4 * q;
            if (al && k + 3 < a.n) {
                v = gld(reinterpret_cast<const float4 *>(h + k));
            } else {
                v.x = k < a.n ? gld(h + k) : 0.f;
                v.y = k + 1 < a.n ? gld(h + k + 1) : 0.f;
                v.z = k + 2 < a.n ? gld(h + k + 2) : 0.f;
                v.w = k + 3 < a.n ? gld(h + k + 3) : 0.f;
            }
        }
        float2 *row = s0 + r * M;
        row[2 * q] = make_float2(v.x, v.y);
        row[2 * q + 1] = make_float2(v.z, v.w);
        row[M / 2 + 2 * q] = make_float2(0.f, 0.f);
        row[M / 2 + 2 * q + 1] = make_float2(0.f, 0.f);
    }
    for (int m = tid; m < M; m += nt) {
        tw[m] = gld(a.twM + m);
        twn[m] = gld(a.twN + m);
    }
    __syncthreads();
    float2 *rows = reinterpret_cast<float2 *>(a.H) + ((size_t)ch * a.P + i0) * M;
    if constexpr (NR == kTickThreads / 64 && M / 4 <= 64) {
        // one row per wave: its stages wait on the wave alone
        const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = tid & 63;
        if (w >= nr) return;
        const float2 *Z = fft_lds<M, false, 64>(s0 + w * M, s1 + w * M, tw);
        for (int k = lane; k < M; k += 64) gst(rows + (size_t)w * M + k, r2c_post<M>(Z, k, twn));
    } else {
        const float2 *Z = fft_lds_n<M, false>(s0, s1, nr, tw);
        for (int e = tid; e < nr * M; e += nt) {
            const int r = e / M, k = e - r * M;
            gst(rows + e, r2c_post<M>(Z + r * M, k, twn));
        }
    }
}

template <int M, int NR>
__global__ __launch_bounds__(256) void k_ir_transform_rows(IrArgs a)
{
    __shared__ float2 s0[NR * M], s1[NR * M], tw[M], twn[M];
    ir_transform_rows_part<M, NR>(a, blockIdx.x, blockIdx.y, s0, s1, tw, twn);
}

// Grouped form of k_ir_transform_rows: ceil(P / NR) blocks per listed channel.
template <int M, int NR>
__global__ __launch_bounds__(256) void k_ir_transform_rows_group(const IrArgs *__restrict__ descs,
                                                                 const int *__restrict__ starts, int n)
{
    __shared__ float2 s0[NR * M], s1[NR * M], tw[M], twn[M];
    const int b = blockIdx.x, j = group_entry(starts, n, b);
    const IrArgs a = descs[j];
    const int local = b - starts[j], chunks = (a.P + NR - 1) / NR;
    ir_transform_rows_part<M, NR>(a, local % chunks, local / chunks, s0, s1, tw, twn);
}

// ---------------------------------------------------------------------------
// TANFFT transforms, one workgroup per channel.  M is the complex FFT size
// (N for complex directions, N/2 for real ones).
// ---------------------------------------------------------------------------
template <int M>
__global__ __launch_bounds__(M / 4 > 64 ? (M / 4 < 1024 ? M / 4 : 1024) : 64) void k_fft(FftArgs a)
{
    __shared__ float2 s0[M > 1 ? M : 2], s1[M > 1 ? M : 2], s2[M > 1 ? M : 2];
    const int c = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const float *in = a.in + (size_t)c * a.stride;
    float *out = a.out + (size_t)c * a.stride;
    const int mode = a.mode;
    if (mode == 0 || mode == 1) {
        for (int n = tid; n < M; n += nt) s0[n] = make_float2(in[2 * n], in[2 * n + 1]);
        __syncthreads();
        const float2 *z = mode == 0 ? fft_lds<M, false>(s0, s1, a.twM) : fft_lds<M, true>(s0, s1, a.twM);
        const float sc = mode == 1 ? 1.0f / (float)M : 1.0f;
        for (int k = tid; k < M; k += nt) {
            out[2 * k] = z[k].x * sc;
            out[2 * k + 1] = z[k].y * sc;
        }
    } else if (mode == 2 || mode == 4) {
        for (int n = tid; n < M; n += nt) s0[n] = make_float2(in[2 * n], in[2 * n + 1]);
        __syncthreads();
        const float2 *Z = fft_lds<M, false>(s0, s1, a.twM);
        const int imoff = M + 8;
        for (int k = tid; k <= M; k += nt) {
            float2 X;
            if (k == 0) X = make_float2(Z[0].x + Z[0].y, 0.f);
            else if (k == M) X = make_float2(Z[0].x - Z[0].y, 0.f);
            else X = r2c_post<M>(Z, k, a.twN);
            if (mode == 2) { out[2 * k] = X.x; out[2 * k + 1] = X.y; }
            else { out[k] = X.x; out[imoff + k] = X.y; }
        }
    } else {
        const int imoff = M + 8;
        // packed spectrum in s2: element 0 = (DC.re, Nyquist.re); imag parts of
        // DC and Nyquist are ignored as a c2r transform does
        for (int k = tid; k < M; k += nt) {
            float2 v;
            if (k == 0) v = mode == 3 ? make_float2(in[0], in[2 * M]) : make_float2(in[0], in[M]);
            else v = mode == 3 ? make_float2(in[2 * k], in[2 * k + 1]) : make_float2(in[k], in[imoff + k]);
            s2[k] = v;
        }
        __syncthreads();
        for (int k = tid; k < M; k += nt) s0[k] = c2r_pre<M>(s2, k, a.twN);
        __syncthreads();
        const float2 *z = fft_lds<M, true>(s0, s1, a.twM);
        const float sc = 1.0f / (2.0f * (float)M);
        for (int n = tid; n < M; n += nt) {
            out[2 * n] = z[n].x * sc;
            out[2 * n + 1] = z[n].y * sc;
        }
    }
}

// ---------------------------------------------------------------------------
// TANFFT above kFftLdsMax complex points: the four-step transform.  With
// M = M1 * M2 and n = n1 * M2 + n2, k = k1 + M1 * k2:
//   Y[k1][n2] = W_M^(n2 k1) * sum_n1 x[n1 M2 + n2] W_M1^(n1 k1)   (columns)
//   X[k1 + M1 k2] = sum_n2 Y[k1][n2] W_M2^(n2 k2)                  (rows)
// Each workgroup transforms T = 4096 / size columns or rows in LDS (64 KB),
// so global reads and writes come in runs of T consecutive elements.  The
// twiddles between the steps are generated in FP64 and rounded once.
// ---------------------------------------------------------------------------
struct Fft4Pass {
    const float2 *in;
    float2 *out;
    long long in_stride, out_stride;  // complex elements between channels
    int logM, logM1;                   // M and the column size M1 (M2 = M / M1)
    int inv;
    float scale;                       // row pass: applied to the outputs
    const float2 *tw;                  // W_M1 (columns) or W_M2 (rows)
};

// exp(-+2 pi i r / 2^logn) rounded once to FP32 (sign + for the inverse)
__device__ __forceinline__ float2 tw_fp64(unsigned long long r, int logn, bool inv)
{
    double s, c;
    sincospi(-2.0 * (double)r / (double)(1ull << logn), &s, &c);
    return make_float2((float)c, (float)(inv ? -s : s));
}

constexpr int kFft4Tile = 4096;  // complex elements a four-step workgroup holds

template <int M1>
__global__ __launch_bounds__(256) void k_fft4_cols(Fft4Pass a)
{
    constexpr int T = kFft4Tile / M1;  // columns per workgroup
    __shared__ float2 s0[kFft4Tile], s1[kFft4Tile];
    const int c = blockIdx.y, tid = threadIdx.x;
    const long long M2 = 1ll << (a.logM - a.logM1);
    const long long n20 = (long long)blockIdx.x * T;
    const float2 *x = a.in + (size_t)c * a.in_stride;
    // column t of the tile as LDS row t; runs of T consecutive n2 per n1
    for (int e = tid; e < kFft4Tile; e += 256) {
        const int t = e % T, n1 = e / T;
        s0[t * M1 + n1] = x[(size_t)n1 * M2 + n20 + t];
    }
    __syncthreads();
    const float2 *z = a.inv ? fft_lds_n<M1, true>(s0, s1, T, a.tw) : fft_lds_n<M1, false>(s0, s1, T, a.tw);
    float2 *y = a.out + (size_t)c * a.out_stride;
    const unsigned long long mask = (1ull << a.logM) - 1;
    for (int e = tid; e < kFft4Tile; e += 256) {
        const int t = e % T, k1 = e / T;
        const long long n2 = n20 + t;
        const float2 w = tw_fp64(((unsigned long long)n2 * (unsigned long long)k1) & mask, a.logM, a.inv != 0);
        y[(size_t)k1 * M2 + n2] = cmul(z[t * M1 + k1], w);
    }
}

template <int M2>
__global__ __launch_bounds__(256) void k_fft4_rows(Fft4Pass a)
{
    constexpr int T = kFft4Tile / M2;  // rows per workgroup
    __shared__ float2 s0[kFft4Tile], s1[kFft4Tile];
    const int c = blockIdx.y, tid = threadIdx.x;
    const long long M1 = 1ll << a.logM1;
    const long long k10 = (long long)blockIdx.x * T;
    const float2 *y = a.in + (size_t)c * a.in_stride + (size_t)k10 * M2;
    for (int e = tid; e < kFft4Tile; e += 256) s0[e] = y[e];
    __syncthreads();
    const float2 *z = a.inv ? fft_lds_n<M2, true>(s0, s1, T, a.tw) : fft_lds_n<M2, false>(s0, s1, T, a.tw);
    float2 *X = a.out + (size_t)c * a.out_stride;
    for (int e = tid; e < kFft4Tile; e += 256) {
        const int t = e % T, k2 = e / T;
        const float2 v = z[t * M2 + k2];
        X[(size_t)k2 * M1 + k10 + t] = make_float2(v.x * a.scale, v.y * a.scale);
    }
}

// Real forward transforms of 2M points from Z = FFT_M(z) (z[n] = x[2n] +
// i x[2n+1]): bins 0..M into the interleaved (mode 2) or planar (mode 4,
// imaginary parts from M + 8) layout, as k_fft writes them.
__global__ __launch_bounds__(256) void k_fft4_r2c_post(const float2 *Z, long long zs, float *out, long long os,
                                                       int logM, int mode)
{
    const int c = blockIdx.y;
    const long long M = 1ll << logM;
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k > M) return;
    const float2 *z = Z + (size_t)c * zs;
    float *o = out + (size_t)c * os;
    float2 X;
    if (k == 0 || k == M) {
        const float2 z0 = z[0];
        X = make_float2(k == 0 ? z0.x + z0.y : z0.x - z0.y, 0.f);
    } else {
        // r2c_post with the W_2M twiddle generated in FP64
        const float2 z0 = z[k], zc = z[M - k];
        const float er = 0.5f * (z0.x + zc.x), ei = 0.5f * (z0.y - zc.y);
        const float dr = z0.x - zc.x, di = z0.y + zc.y;
        const float orr = 0.5f * di, oi = -0.5f * dr;
        const float2 w = tw_fp64((unsigned long long)k, logM + 1, false);
        X = make_float2(er + fmaf(w.x, orr, -(w.y * oi)), ei + fmaf(w.x, oi, w.y * orr));
    }
    if (mode == 2) {
        o[2 * k] = X.x;
        o[2 * k + 1] = X.y;
    } else {
        o[k] = X.x;
        o[M + 8 + k] = X.y;
    }
}

// Real inverse transforms: the packed M-point input of c2r_pre from the
// interleaved (mode 3) or planar (mode 5) spectrum, as k_fft reads it.
__device__ __forceinline__ float2 c2r_bin(const float *in, long long M, long long k, int mode)
{
    if (k == 0) return mode == 3 ? make_float2(in[0], in[2 * M]) : make_float2(in[0], in[M]);
    return mode == 3 ? make_float2(in[2 * k], in[2 * k + 1]) : make_float2(in[k], in[M + 8 + k]);
}

__global__ __launch_bounds__(256) void k_fft4_c2r_pre(const float *in, long long is, float2 *A, long long as,
                                                      int logM, int mode)
{
    const int c = blockIdx.y;
    const long long M = 1ll << logM;
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= M) return;
    const float *x = in + (size_t)c * is;
    const float2 a = c2r_bin(x, M, k, mode);
    float2 r;
    if (k == 0) {
        r = make_float2(a.x + a.y, a.x - a.y);
    } else {
        const float2 b = c2r_bin(x, M, M - k, mode);
        const float er = a.x + b.x, ei = a.y - b.y;
        const float dr = a.x - b.x, di = a.y + b.y;
        const float2 w = tw_fp64((unsigned long long)k, logM + 1, false);
        const float orr = fmaf(dr, w.x, di * w.y), oi = fmaf(di, w.x, -(dr * w.y));
        r = make_float2(er - oi, ei + orr);
    }
    A[(size_t)c * as + k] = r;
}

// ---------------------------------------------------------------------------
// Time-domain convolution (see TdArgs), 256 threads, window and filters in LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_td(TdArgs a)
{
    HIP_DYNAMIC_SHARED(float, td_smem)
    const int c = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    if (!a.run[c]) return;
    const int len = a.len, n = a.n, mask = len - 1;
    int f0 = 0, f1 = len;
    if (a.nz) {
        f0 = max(0, min(len, a.nz[2 * c]));
        f1 = max(f0, min(len, a.nz[2 * c + 1]));
    }
    // window w[t] = sample (pos - (len - 1) + t) for t < len - 1 + n: the old
    // history, then the n new samples
    float *w = td_smem;                     // [len - 1 + n]
    float *ha = td_smem + (len - 1 + a.B);  // [len]
    float *hb = ha + len;                   // [len] (crossfade)
    const float *hist = a.hist + (size_t)c * len;
    const float *xi = a.xin + (size_t)c * a.in_stride;
    for (int t = tid; t < len - 1; t += nt) w[t] = hist[(a.pos - (len - 1) + t) & mask];
    for (int j = tid; j < n; j += nt) w[len - 1 + j] = xi[j];
    // the taps the sums use, [f0, f1), only: a caller's IR array may hold no more
    const float *Ha = a.Ha + (size_t)c * a.h_stride;
    for (int k = f0 + tid; k < f1; k += nt) ha[k] = Ha[k];
    if (a.Hb) {
        const float *Hb = a.Hb + (size_t)c * a.h_stride;
        for (int k = f0 + tid; k < f1; k += nt) hb[k] = Hb[k];
    }
    __syncthreads();
    float *yo = a.yout + (size_t)c * a.out_stride;
    const float step = (float)(1.0 / (double)a.B);
    // y[j] = sum_k w[len - 1 + j - k] h[k] in 8 interleaved partial sums
    // (taps k = f0 + 8q + r go to sum r): a long FP32 sum in one accumulator
    // loses about 9 dB of SNR at 4096 taps
    auto dot = [&](const float *wj, const float *h) {
        float p[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        int k = f0;
        for (; k + 8 <= f1; k += 8) {
#pragma unroll
            for (int r = 0; r < 8; r++) p[r] += wj[-(k + r)] * h[k + r];
        }
        for (int r = 0; k < f1; k++, r++) p[r] += wj[-k] * h[k];
        return ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
    };
    for (int j = tid; j < n; j += nt) {
        const float *wj = w + (len - 1 + j);
        float ya = dot(wj, ha);
        if (a.Hb) {
            const float yb = dot(wj, hb);
            ya = yb * (step * (float)j) + ya * (step * (float)(a.B - j));
        }
        yo[j] = ya;
    }
    // the new samples enter the history (this workgroup alone owns channel c)
    float *hw = a.hist + (size_t)c * len;
    for (int j = tid; j < n; j += nt) hw[(a.pos + j) & mask] = xi[j];
}

__global__ void k_td_load(const float *__restrict__ ir, long long stride, long long n, const int *__restrict__ chan,
                          float *__restrict__ H, int len)
{
    const int ch = chan ? chan[blockIdx.y] : (int)blockIdx.y;
    const float *src = ir + (size_t)ch * stride;
    float *dst = H + (size_t)ch * len;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < len; k += gridDim.x * blockDim.x)
        dst[k] = k < n ? src[k] : 0.f;
}

// ---------------------------------------------------------------------------
// Block DFT on the matrix cores: a batched real-to-complex DFT of frames of
// n samples as the GEMM Y[c][col] = sum_t x[c][t] F[t][col], F[t][2k] =
// cos(2 pi k t / n), F[t][2k+1] = -sin(2 pi k t / n), k = 0..n/2.  bf16
// operands with FP32 accumulation; with split = 3 each operand is split into
// bf16 high and low parts and three products (hi hi, hi lo, lo hi) are
// accumulated, which recovers about 16 more bits of the products.
// v_mfma_f32_16x16x32_bf16 lane layout: A (16 x 32): lane l holds row l % 16,
// k = 8 (l / 16) + e; B (32 x 16): column l % 16, same k; D (16 x 16): column
// l % 16, rows 4 (l / 16) + r.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int kDftWaves = 4;  // waves per workgroup, each over a quarter of the samples

__global__ __launch_bounds__(64 * kDftWaves) void k_dft_mfma(const float *__restrict__ x, long long in_stride,
                                                             float *__restrict__ y, long long out_stride, int C,
                                                             int n, int cols, const bf16x8_t *__restrict__ fh,
                                                             const bf16x8_t *__restrict__ fl, int split)
{
    // a 32-channel x 32-column tile as 2 x 2 MFMA tiles per wave, so each
    // k-step's operand fragments feed four (or twelve) products; the waves
    // split the sum over t and their partial tiles are added in LDS in wave
    // order (the result does not depend on scheduling)
    __shared__ float part[kDftWaves][16][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int tiles_c = (cols + 31) / 32;
    const int c0 = (blockIdx.x / tiles_c) * 32, col0 = (blockIdx.x % tiles_c) * 32;
    if (c0 >= C) return;  // uniform over the workgroup
    const int i = lane & 15, kq = lane >> 4;
    const float *xr[2];
    int colr[2];
#pragma unroll
    for (int u = 0; u < 2; u++) {
        xr[u] = x + (size_t)min(c0 + 16 * u + i, C - 1) * in_stride;  // rows past C compute a copy, not stored
        colr[u] = min(col0 + 16 * u + i, cols - 1);
    }
    f32x4_t acc[2][2];
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
        for (int v = 0; v < 2; v++) acc[u][v] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
    const int kv = n / 8;  // bf16x8 vectors per table row
    const int steps = n / 32, per = (steps + kDftWaves - 1) / kDftWaves;
    const int s0 = w * per, s1 = min(steps, s0 + per);
    for (int st = s0; st < s1; st++) {
        const int kb = 32 * st + 8 * kq;
        bf16x8_t ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const float4 p0 = *reinterpret_cast<const float4 *>(xr[u] + kb);
            const float4 p1 = *reinterpret_cast<const float4 *>(xr[u] + kb + 4);
            const float v8[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const __bf16 h = (__bf16)v8[e];
                ah[u][e] = h;
                al[u][e] = (__bf16)(v8[e] - (float)h);
            }
        }
#pragma unroll
        for (int v = 0; v < 2; v++) {
            bh[v] = fh[(size_t)colr[v] * kv + kb / 8];
            if (split == 3) bl[v] = fl[(size_t)colr[v] * kv + kb / 8];
        }
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int v = 0; v < 2; v++) {
                acc[u][v] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[u], bh[v], acc[u][v], 0, 0, 0);
                if (split == 3) {
                    acc[u][v] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[u], bl[v], acc[u][v], 0, 0, 0);
                    acc[u][v] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[u], bh[v], acc[u][v], 0, 0, 0);
                }
            }
    }
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
        for (int v = 0; v < 2; v++)
#pragma unroll
            for (int r = 0; r < 4; r++) part[w][(u * 2 + v) * 4 + r][lane] = acc[u][v][r];
    __syncthreads();
    // 1,024 outputs, four per thread: element e = (u, v, r) of lane l
    for (int o = threadIdx.x; o < 16 * 64; o += 64 * kDftWaves) {
        const int e = o >> 6, l = o & 63;
        float sum = part[0][e][l];
#pragma unroll
        for (int ww = 1; ww < kDftWaves; ww++) sum += part[ww][e][l];
        const int u = e >> 3, v = (e >> 2) & 1, r = e & 3;
        const int row = c0 + 16 * u + 4 * (l >> 4) + r, col = col0 + 16 * v + (l & 15);
        if (row < C && col < cols) y[(size_t)row * out_stride + col] = sum;
    }
}

// ---------------------------------------------------------------------------
// Output chain: mixer and sample converter (grid-stride, one sample per lane).
// ---------------------------------------------------------------------------
// Sum of the channels in channel order (the reference's order), channel c at
// ins[c] (device pointer list) or at in + c * stride.
__global__ void k_mix(const float *const *__restrict__ ins, const float *__restrict__ in, long long stride,
                      int channels, long long n, float *__restrict__ out)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float acc = 0.f;
        for (int c = 0; c < channels; c++) {
            const float v = ins ? ins[c][i] : in[(size_t)c * stride + i];
            acc = c == 0 ? v : acc + v;
        }
        out[i] = acc;
    }
}

// IT/OT: 0 = float, 1 = short.  float -> short scales by 32767 * gain,
// clamps to [-32768, 32767] (setting *flag) and converts toward zero;
// short -> float scales by gain / 32767.
template <int IT, int OT>
__global__ void k_convert(const void *__restrict__ in, long long in_step, long long in_off, void *__restrict__ out,
                          long long out_step, long long out_off, long long n, float gain, int *__restrict__ flag)
{
    bool clip = false;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const long long si = i * in_step + in_off, so = i * out_step + out_off;
        if constexpr (IT == 0 && OT == 1) {
            float f = static_cast<const float *>(in)[si] * (32767.0f * gain);
            if (f > 32767.0f) { f = 32767.0f; clip = true; }
            if (f < -32768.0f) { f = -32768.0f; clip = true; }
            static_cast<short *>(out)[so] = (short)(int)f;
        } else if constexpr (IT == 1 && OT == 0) {
            static_cast<float *>(out)[so] = (float)static_cast<const short *>(in)[si] * (gain / 32767.0f);
        } else if constexpr (IT == 0) {
            static_cast<float *>(out)[so] = static_cast<const float *>(in)[si];
        } else {
            static_cast<short *>(out)[so] = static_cast<const short *>(in)[si];
        }
    }
    if (clip && flag) *flag = 1;
}

__global__ void k_zero(float4 *p, size_t n4)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline int fft_threads(int M)
{
    int t = M / 4;
    if (t < 64) t = 64;
    if (t > 1024) t = 1024;
    return t;
}

#define TANX_SWITCH_B(B, CALL)          \
    switch (B) {                        \
    case 64: { constexpr int M = 64; CALL; } break;     \
    case 128: { constexpr int M = 128; CALL; } break;   \
    case 256: { constexpr int M = 256; CALL; } break;   \
    case 512: { constexpr int M = 512; CALL; } break;   \
    case 1024: { constexpr int M = 1024; CALL; } break; \
    default: return hipErrorInvalidValue;               \
    }

// With ev0/ev1, the launch itself records the kernel's start and end (the
// dispatch's own timestamps), so event timing measures the kernel alone.
template <int HT>
static hipError_t launch_tick_t(int B, int tail_nf, const TickArgs &a, hipStream_t s, unsigned blocks,
                                hipEvent_t ev0, hipEvent_t ev1)
{
    const int tk = a.nchunk == 0 ? 0 : (is_pass(a) ? 2 : 1);
    if (tk == 0) {
        TANX_SWITCH_B(B, hipExtLaunchKernelGGL((k_tick<M, 1, HT, 0, 4>), dim3(blocks), dim3(kTickThreads), 0, s, ev0, ev1, 0, a));
    } else if (tk == 1) {
        if (tail_nf == 2) {
            TANX_SWITCH_B(B, hipExtLaunchKernelGGL((k_tick<M, 2, HT, 1, 4>), dim3(blocks), dim3(kTickThreads), 0, s, ev0, ev1, 0, a));
        } else {
            TANX_SWITCH_B(B, hipExtLaunchKernelGGL((k_tick<M, 1, HT, 1, 4>), dim3(blocks), dim3(kTickThreads), 0, s, ev0, ev1, 0, a));
        }
    } else if (pass_km(a.kp) == 4) {
        if (tail_nf == 2) {
            TANX_SWITCH_B(B, hipExtLaunchKernelGGL((k_tick<M, 2, HT, 2, 4>), dim3(blocks), dim3(kTickThreads), 0, s, ev0, ev1, 0, a));
        } else {
            TANX_SWITCH_B(B, hipExtLaunchKernelGGL((k_tick<M, 1, HT, 2, 4>), dim3(blocks), dim3(kTickThreads), 0, s, ev0, ev1, 0, a));
        }
    } else if (pass_km(a.kp) == 8) {
        if (tail_nf == 2) {
            TANX_SWITCH_B(B, hipExtLaunchKernelGGL((k_tick<M, 2, HT, 2, 8>), dim3(blocks), dim3(kTickThreads), 0, s, ev0, ev1, 0, a));
        } else {
            TANX_SWITCH_B(B, hipExtLaunchKernelGGL((k_tick<M, 1, HT, 2, 8>), dim3(blocks), dim3(kTickThreads), 0, s, ev0, ev1, 0, a));
        }
    } else if (pass_km(a.kp) == 12) {
        if (tail_nf == 2) {
            TANX_SWITCH_B(B, hipExtLaunchKernelGGL((k_tick<M, 2, HT, 2, 12>), dim3(blocks), dim3(kTickThreads), 0, s, ev0, ev1, 0, a));
        } else {
            TANX_SWITCH_B(B, hipExtLaunchKernelGGL((k_tick<M, 1, HT, 2, 12>), dim3(blocks), dim3(kTickThreads), 0, s, ev0, ev1, 0, a));
        }
    } else {
        if (tail_nf == 2) {
            TANX_SWITCH_B(B, hipExtLaunchKernelGGL((k_tick<M, 2, HT, 2, 16>), dim3(blocks), dim3(kTickThreads), 0, s, ev0, ev1, 0, a));
        } else {
            TANX_SWITCH_B(B, hipExtLaunchKernelGGL((k_tick<M, 1, HT, 2, 16>), dim3(blocks), dim3(kTickThreads), 0, s, ev0, ev1, 0, a));
        }
    }
    return hipGetLastError();
}

hipError_t launch_tick(int B, int tail_nf, const TickArgs &a, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    const unsigned blocks = (unsigned)(a.nhead + a.C * a.nchunk + a.C * a.nchB);
    if (blocks == 0) return hipSuccess;
    switch (a.irtype) {
    case IR_F32: return launch_tick_t<IR_F32>(B, tail_nf, a, s, blocks, ev0, ev1);
    case IR_F16: return launch_tick_t<IR_F16>(B, tail_nf, a, s, blocks, ev0, ev1);
    case IR_BF16: return launch_tick_t<IR_BF16>(B, tail_nf, a, s, blocks, ev0, ev1);
    default: return hipErrorInvalidValue;
    }
}

template <int HT>
static hipError_t launch_desc_t(int B, const TickArgs *d, int k, unsigned blocks, hipStream_t s)
{
    TANX_SWITCH_B(B, hipLaunchKernelGGL((k_tick_desc<M, HT>), dim3(blocks), dim3(kTickThreads), 0, s, d, k));
    return hipGetLastError();
}

hipError_t launch_tick_desc(int B, int irtype, const TickArgs *d_descs, int k, unsigned blocks, hipStream_t s)
{
    switch (irtype) {
    case IR_F32: return launch_desc_t<IR_F32>(B, d_descs, k, blocks, s);
    case IR_F16: return launch_desc_t<IR_F16>(B, d_descs, k, blocks, s);
    case IR_BF16: return launch_desc_t<IR_BF16>(B, d_descs, k, blocks, s);
    default: return hipErrorInvalidValue;
    }
}

template <int HT, int M>
static hipError_t launch_group_m(int roles, int km, const TickArgs *d, const int *st, int nobj, unsigned blocks,
                                 hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    const dim3 g(blocks), t(kTickThreads);
    if (km != 4 && km != 8 && km != 12 && km != 16) return hipErrorInvalidValue;
#define TANX_GROUP_KM(R)                                                                                    \
    if (km == 4) hipExtLaunchKernelGGL((k_group_tick<HT, M, R, 4>), g, t, 0, s, ev0, ev1, 0, d, st, nobj);       \
    else if (km == 8) hipExtLaunchKernelGGL((k_group_tick<HT, M, R, 8>), g, t, 0, s, ev0, ev1, 0, d, st, nobj);  \
    else if (km == 12) hipExtLaunchKernelGGL((k_group_tick<HT, M, R, 12>), g, t, 0, s, ev0, ev1, 0, d, st, nobj); \
    else hipExtLaunchKernelGGL((k_group_tick<HT, M, R, 16>), g, t, 0, s, ev0, ev1, 0, d, st, nobj)
    switch (roles) {
    case kRolesGeneral:
        hipExtLaunchKernelGGL((k_group_tick<HT, M, kRolesGeneral, 4>), g, t, 0, s, ev0, ev1, 0, d, st, nobj);
        break;
    case kRolesLean: TANX_GROUP_KM(kRolesLean); break;
    case kRolesXfade: TANX_GROUP_KM(kRolesXfade); break;
    case kRolesHeads: hipExtLaunchKernelGGL((k_group_heads<HT, M>), g, t, 0, s, ev0, ev1, 0, d, st, nobj); break;
    default: return hipErrorInvalidValue;
    }
#undef TANX_GROUP_KM
    return hipSuccess;
}

template <int HT>
static hipError_t launch_group_t(int B, int roles, int km, const TickArgs *d, const int *st, int nobj,
                                 unsigned blocks, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    switch (B) {
    case 64: return launch_group_m<HT, 64>(roles, km, d, st, nobj, blocks, s, ev0, ev1);
    case 128: return launch_group_m<HT, 128>(roles, km, d, st, nobj, blocks, s, ev0, ev1);
    case 256: return launch_group_m<HT, 256>(roles, km, d, st, nobj, blocks, s, ev0, ev1);
    case 512: return launch_group_m<HT, 512>(roles, km, d, st, nobj, blocks, s, ev0, ev1);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_group_tick(int irtype, int B, int roles, int km, const TickArgs *d_descs, const int *d_starts,
                             int nobj, unsigned blocks, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    if (blocks == 0) return hipSuccess;
    hipError_t e;
    switch (irtype) {
    case IR_F32: e = launch_group_t<IR_F32>(B, roles, km, d_descs, d_starts, nobj, blocks, s, ev0, ev1); break;
    case IR_F16: e = launch_group_t<IR_F16>(B, roles, km, d_descs, d_starts, nobj, blocks, s, ev0, ev1); break;
    case IR_BF16: e = launch_group_t<IR_BF16>(B, roles, km, d_descs, d_starts, nobj, blocks, s, ev0, ev1); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

// Resident workgroups of one kernel on the current device (cached per kernel)
static unsigned resident_blocks(const void *kernel)
{
    static std::mutex mtx;
    static std::map<std::pair<int, const void *>, unsigned> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lk(mtx);
    auto it = cache.find({dev, kernel});
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kTickThreads, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    const unsigned n = (unsigned)std::max(1, per_cu) * (unsigned)std::max(1, cus);
    cache[{dev, kernel}] = n;
    return n;
}

template <int HT, int M, int ROLES, int KM>
static hipError_t launch_queue_k(const TickArgs *d, const int *st, int nobj, int total, int *next, hipStream_t s,
                                 hipEvent_t ev0, hipEvent_t ev1)
{
    const void *k = reinterpret_cast<const void *>(&k_group_queue<HT, M, ROLES, KM>);
    const unsigned res = resident_blocks(k);
    if (res == 0) return hipErrorInvalidValue;
    const unsigned grid = std::min(res, (unsigned)total);
    hipExtLaunchKernelGGL((k_group_queue<HT, M, ROLES, KM>), dim3(grid), dim3(kTickThreads), 0, s, ev0, ev1, 0, d, st,
                          nobj, total, next);
    return hipSuccess;
}

template <int HT, int M>
static hipError_t launch_queue_m(int roles, int km, const TickArgs *d, const int *st, int nobj, int total, int *next,
                                 hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    if (km != 4 && km != 8 && km != 12 && km != 16) return hipErrorInvalidValue;
#define TANX_QUEUE_KM(R)                                                                  \
    (km == 4 ? launch_queue_k<HT, M, R, 4>(d, st, nobj, total, next, s, ev0, ev1)        \
     : km == 8 ? launch_queue_k<HT, M, R, 8>(d, st, nobj, total, next, s, ev0, ev1)      \
     : km == 12 ? launch_queue_k<HT, M, R, 12>(d, st, nobj, total, next, s, ev0, ev1)    \
                : launch_queue_k<HT, M, R, 16>(d, st, nobj, total, next, s, ev0, ev1))
    hipError_t e;
    switch (roles) {
    case kRolesGeneral: e = launch_queue_k<HT, M, kRolesGeneral, 4>(d, st, nobj, total, next, s, ev0, ev1); break;
    case kRolesLean: e = TANX_QUEUE_KM(kRolesLean); break;
    case kRolesXfade: e = TANX_QUEUE_KM(kRolesXfade); break;
    case kRolesHeads: e = launch_queue_k<HT, M, kRolesHeads, 4>(d, st, nobj, total, next, s, ev0, ev1); break;
    default: return hipErrorInvalidValue;
    }
#undef TANX_QUEUE_KM
    return e;
}

template <int HT>
static hipError_t launch_queue_t(int B, int roles, int km, const TickArgs *d, const int *st, int nobj, int total,
                                 int *next, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    switch (B) {
    case 64: return launch_queue_m<HT, 64>(roles, km, d, st, nobj, total, next, s, ev0, ev1);
    case 128: return launch_queue_m<HT, 128>(roles, km, d, st, nobj, total, next, s, ev0, ev1);
    case 256: return launch_queue_m<HT, 256>(roles, km, d, st, nobj, total, next, s, ev0, ev1);
    case 512: return launch_queue_m<HT, 512>(roles, km, d, st, nobj, total, next, s, ev0, ev1);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_group_queue(int irtype, int B, int roles, int km, const TickArgs *d_descs, const int *d_starts,
                              int nobj, int total, int *next, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    if (total <= 0) return hipSuccess;
    hipError_t e;
    switch (irtype) {
    case IR_F32: e = launch_queue_t<IR_F32>(B, roles, km, d_descs, d_starts, nobj, total, next, s, ev0, ev1); break;
    case IR_F16: e = launch_queue_t<IR_F16>(B, roles, km, d_descs, d_starts, nobj, total, next, s, ev0, ev1); break;
    case IR_BF16: e = launch_queue_t<IR_BF16>(B, roles, km, d_descs, d_starts, nobj, total, next, s, ev0, ev1); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

template <int HT>
static hipError_t launch_heads_t(int B, const HeadBatchArgs &h, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    const dim3 g((unsigned)h.C), t(kTickThreads);
#define TANX_HEADS(MM)                                                                            \
    if (pass_km(h.kp) == 4) hipExtLaunchKernelGGL((k_heads<MM, HT, kHeadBatch + 3>), g, t, 0, s, ev0, ev1, 0, h); \
    else if (pass_km(h.kp) == 8) hipExtLaunchKernelGGL((k_heads<MM, HT, kHeadBatch + 7>), g, t, 0, s, ev0, ev1, 0, h); \
    else if (pass_km(h.kp) == 12) hipExtLaunchKernelGGL((k_heads<MM, HT, kHeadBatch + 11>), g, t, 0, s, ev0, ev1, 0, h); \
    else hipExtLaunchKernelGGL((k_heads<MM, HT, kHeadBatch + 15>), g, t, 0, s, ev0, ev1, 0, h)
    switch (B) {
    case 64: TANX_HEADS(64); break;
    case 128: TANX_HEADS(128); break;
    case 256: TANX_HEADS(256); break;
    default: return hipErrorInvalidValue;
    }
#undef TANX_HEADS
    return hipGetLastError();
}

template <int HT>
static hipError_t launch_window_t(int B, int tail_nf, const HeadBatchArgs &h, const TickArgs &a, hipStream_t s,
                                  hipEvent_t ev0, hipEvent_t ev1)
{
    const dim3 grid((unsigned)(h.C + a.C * a.nchunk));
#define TANX_WINDOW_K(MM, KK)                                                                                     \
    if (tail_nf == 2) hipExtLaunchKernelGGL((k_window<MM, 2, HT, KK>), grid, dim3(kTickThreads), 0, s, ev0, ev1, 0, h, a); \
    else hipExtLaunchKernelGGL((k_window<MM, 1, HT, KK>), grid, dim3(kTickThreads), 0, s, ev0, ev1, 0, h, a)
#define TANX_WINDOW(MM)                              \
    if (pass_km(a.kp) == 4) { TANX_WINDOW_K(MM, 4); }      \
    else if (pass_km(a.kp) == 8) { TANX_WINDOW_K(MM, 8); } \
    else if (pass_km(a.kp) == 12) { TANX_WINDOW_K(MM, 12); } \
    else { TANX_WINDOW_K(MM, 16); }
    switch (B) {
    case 64: TANX_WINDOW(64); break;
    case 128: TANX_WINDOW(128); break;
    case 256: TANX_WINDOW(256); break;
    default: return hipErrorInvalidValue;
    }
#undef TANX_WINDOW
#undef TANX_WINDOW_K
    return hipGetLastError();
}

hipError_t launch_window(int B, int tail_nf, const HeadBatchArgs &h, const TickArgs &a, hipStream_t s,
                         hipEvent_t ev0, hipEvent_t ev1)
{
    if (h.n < 1 || h.n > kHeadBatch || B > kHeadBatchMaxB || a.nchunk < 1 || !is_pass(a) || a.C != h.C ||
        a.kp != h.kp || a.kn > pass_km(a.kp))
        return hipErrorInvalidValue;
    switch (a.irtype) {
    case IR_F32: return launch_window_t<IR_F32>(B, tail_nf, h, a, s, ev0, ev1);
    case IR_F16: return launch_window_t<IR_F16>(B, tail_nf, h, a, s, ev0, ev1);
    case IR_BF16: return launch_window_t<IR_BF16>(B, tail_nf, h, a, s, ev0, ev1);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_heads(int B, int irtype, const HeadBatchArgs &h, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    if (h.n < 1 || h.n > kHeadBatch || B > kHeadBatchMaxB) return hipErrorInvalidValue;
    switch (irtype) {
    case IR_F32: return launch_heads_t<IR_F32>(B, h, s, ev0, ev1);
    case IR_F16: return launch_heads_t<IR_F16>(B, h, s, ev0, ev1);
    case IR_BF16: return launch_heads_t<IR_BF16>(B, h, s, ev0, ev1);
    default: return hipErrorInvalidValue;
    }
}

// partition sizes of uniform blocks (64..1024) and of frames (128..4096)
#define TANX_SWITCH_PART(B, CALL)                        \
    switch (B) {                                         \
    case 64: { constexpr int M = 64; CALL; } break;      \
    case 128: { constexpr int M = 128; CALL; } break;    \
    case 256: { constexpr int M = 256; CALL; } break;    \
    case 512: { constexpr int M = 512; CALL; } break;    \
    case 1024: { constexpr int M = 1024; CALL; } break;  \
    case 2048: { constexpr int M = 2048; CALL; } break;  \
    case 4096: { constexpr int M = 4096; CALL; } break;  \
    default: return hipErrorInvalidValue;                \
    }

int ir_transform_blocks(int B, int irtype, int P, int nchan)
{
    if (irtype == IR_F32 && B <= 256) return (P + 3) / 4 * nchan;
    if (irtype == IR_F32 && B == 512) return (P + 1) / 2 * nchan;
    return P * nchan;
}

hipError_t launch_ir_transform_group(int B, int irtype, const IrArgs *d_descs, const int *d_starts, int n,
                                     unsigned blocks, hipStream_t s)
{
    if (n == 0 || blocks == 0) return hipSuccess;
    const dim3 g(blocks);
    switch (irtype) {
    case IR_F32:
        if (B <= 256) {
            switch (B) {
            case 64: hipLaunchKernelGGL((k_ir_transform_rows_group<64, 4>), g, dim3(256), 0, s, d_descs, d_starts, n); break;
            case 128: hipLaunchKernelGGL((k_ir_transform_rows_group<128, 4>), g, dim3(256), 0, s, d_descs, d_starts, n); break;
            case 256: hipLaunchKernelGGL((k_ir_transform_rows_group<256, 4>), g, dim3(256), 0, s, d_descs, d_starts, n); break;
            default: return hipErrorInvalidValue;
            }
        } else if (B == 512) {
            hipLaunchKernelGGL((k_ir_transform_rows_group<512, 2>), g, dim3(256), 0, s, d_descs, d_starts, n);
        } else {
            TANX_SWITCH_PART(B, hipLaunchKernelGGL((k_ir_transform_group<M, IR_F32>), g, dim3(fft_threads(M)), 0, s,
                                                   d_descs, d_starts, n));
        }
        break;
    case IR_F16:
        TANX_SWITCH_PART(B, hipLaunchKernelGGL((k_ir_transform_group<M, IR_F16>), g, dim3(fft_threads(M)), 0, s,
                                               d_descs, d_starts, n));
        break;
    case IR_BF16:
        TANX_SWITCH_PART(B, hipLaunchKernelGGL((k_ir_transform_group<M, IR_BF16>), g, dim3(fft_threads(M)), 0, s,
                                               d_descs, d_starts, n));
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_ir_transform(int B, int nchan, const IrArgs &a, hipStream_t s)
{
    if (a.P == 0 || nchan == 0) return hipSuccess;
    switch (a.irtype) {
    case IR_F32:
        // short partitions: several rows per workgroup
        if (B <= 256) {
            const dim3 g((unsigned)((a.P + 3) / 4), (unsigned)nchan);
            switch (B) {
            case 64: hipLaunchKernelGGL((k_ir_transform_rows<64, 4>), g, dim3(256), 0, s, a); break;
            case 128: hipLaunchKernelGGL((k_ir_transform_rows<128, 4>), g, dim3(256), 0, s, a); break;
            case 256: hipLaunchKernelGGL((k_ir_transform_rows<256, 4>), g, dim3(256), 0, s, a); break;
            default: return hipErrorInvalidValue;
            }
        } else if (B == 512) {
            hipLaunchKernelGGL((k_ir_transform_rows<512, 2>), dim3((unsigned)((a.P + 1) / 2), (unsigned)nchan), dim3(256),
                               0, s, a);
        } else {
            TANX_SWITCH_PART(B, hipLaunchKernelGGL((k_ir_transform<M, IR_F32>), dim3(a.P, nchan), dim3(fft_threads(M)), 0, s, a));
        }
        break;
    case IR_F16:
        TANX_SWITCH_PART(B, hipLaunchKernelGGL((k_ir_transform<M, IR_F16>), dim3(a.P, nchan), dim3(fft_threads(M)), 0, s, a));
        break;
    case IR_BF16:
        TANX_SWITCH_PART(B, hipLaunchKernelGGL((k_ir_transform<M, IR_BF16>), dim3(a.P, nchan), dim3(fft_threads(M)), 0, s, a));
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

#define TANX_SWITCH_FRAME(N, CALL)                       \
    switch (N) {                                         \
    case 128: { constexpr int F = 128; CALL; } break;    \
    case 256: { constexpr int F = 256; CALL; } break;    \
    case 512: { constexpr int F = 512; CALL; } break;    \
    case 1024: { constexpr int F = 1024; CALL; } break;  \
    case 2048: { constexpr int F = 2048; CALL; } break;  \
    case 4096: { constexpr int F = 4096; CALL; } break;  \
    default: return hipErrorInvalidValue;                \
    }

hipError_t launch_frame(int irtype, const FrameArgs &a, hipStream_t s, hipEvent_t ev1)
{
    if (a.C == 0) return hipSuccess;
    switch (irtype) {
    case IR_F32: TANX_SWITCH_FRAME(a.N, hipExtLaunchKernelGGL((k_frame<F, IR_F32>), dim3(a.C), dim3(frame_threads<F>()), 0, s, nullptr, ev1, 0, a)); break;
    case IR_F16: TANX_SWITCH_FRAME(a.N, hipExtLaunchKernelGGL((k_frame<F, IR_F16>), dim3(a.C), dim3(frame_threads<F>()), 0, s, nullptr, ev1, 0, a)); break;
    case IR_BF16: TANX_SWITCH_FRAME(a.N, hipExtLaunchKernelGGL((k_frame<F, IR_BF16>), dim3(a.C), dim3(frame_threads<F>()), 0, s, nullptr, ev1, 0, a)); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_group_frame(int irtype, const FrameArgs *d_descs, const int *d_starts, int nobj, unsigned blocks,
                              hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    if (blocks == 0 || nobj == 0) return hipSuccess;
    const dim3 g(blocks), t(kGroupFrameThreads);
    switch (irtype) {
    case IR_F32: hipExtLaunchKernelGGL((k_group_frame<IR_F32>), g, t, 0, s, ev0, ev1, 0, d_descs, d_starts, nobj); break;
    case IR_F16: hipExtLaunchKernelGGL((k_group_frame<IR_F16>), g, t, 0, s, ev0, ev1, 0, d_descs, d_starts, nobj); break;
    case IR_BF16: hipExtLaunchKernelGGL((k_group_frame<IR_BF16>), g, t, 0, s, ev0, ev1, 0, d_descs, d_starts, nobj); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

static hipError_t launch_fft4_pass(bool rows, int size, int channels, const Fft4Pass &p, hipStream_t s)
{
    const unsigned tiles = (unsigned)((1ll << p.logM) / kFft4Tile);
    const dim3 g(tiles, (unsigned)channels), t(256);
    switch (size) {
#define TANX_FFT4_CASE(V)                                                  \
    case V:                                                                \
        if (rows) hipLaunchKernelGGL(k_fft4_rows<V>, g, t, 0, s, p);       \
        else hipLaunchKernelGGL(k_fft4_cols<V>, g, t, 0, s, p);            \
        break;
        TANX_FFT4_CASE(64) TANX_FFT4_CASE(128) TANX_FFT4_CASE(256) TANX_FFT4_CASE(512) TANX_FFT4_CASE(1024)
        TANX_FFT4_CASE(2048)
#undef TANX_FFT4_CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// The four-step transform of M complex points per channel (M > kFftLdsMax):
// complex directions from/to the caller's buffers; real ones through the
// r2c/c2r kernels.  Scratch: two arrays of M complex per channel.
static hipError_t launch_fft4(int channels, const FftArgs &a, hipStream_t s)
{
    const bool real = a.mode >= 2;
    const int logM = real ? a.log2len - 1 : a.log2len;
    const long long M = 1ll << logM;
    const int logM1 = logM / 2;  // column size <= row size
    if (M <= kFftLdsMax || logM > kFftLog2Max || !a.scratch || !a.tw1 || !a.tw2) return hipErrorInvalidValue;
    const long long cs = a.stride / 2;  // the caller's channel stride in complex elements
    float2 *A = a.scratch, *Bs = a.scratch + (size_t)channels * M;
    const bool inv = a.mode == 1 || a.mode == 3 || a.mode == 5;
    Fft4Pass cols{}, rows{};
    cols.logM = rows.logM = logM;
    cols.logM1 = rows.logM1 = logM1;
    cols.inv = rows.inv = inv ? 1 : 0;
    cols.tw = a.tw1;
    rows.tw = a.tw2;
    rows.scale = a.mode == 1 ? 1.0f / (float)M : (real && inv) ? 1.0f / (2.0f * (float)M) : 1.0f;
    const unsigned bins = (unsigned)((M + 1 + 255) / 256);
    hipError_t e;
    if (!real) {
        cols.in = reinterpret_cast<const float2 *>(a.in);
        cols.in_stride = cs;
        cols.out = A;
        cols.out_stride = M;
        rows.in = A;
        rows.in_stride = M;
        rows.out = reinterpret_cast<float2 *>(a.out);
        rows.out_stride = cs;
    } else if (!inv) {
        cols.in = reinterpret_cast<const float2 *>(a.in);
        cols.in_stride = cs;
        cols.out = A;
        cols.out_stride = M;
        rows.in = A;
        rows.in_stride = M;
        rows.out = Bs;
        rows.out_stride = M;
    } else {
        hipLaunchKernelGGL(k_fft4_c2r_pre, dim3(bins, (unsigned)channels), dim3(256), 0, s, a.in, a.stride, A, M,
                           logM, a.mode);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        cols.in = A;
        cols.in_stride = M;
        cols.out = Bs;
        cols.out_stride = M;
        rows.in = Bs;
        rows.in_stride = M;
        rows.out = reinterpret_cast<float2 *>(a.out);
        rows.out_stride = cs;
    }
    if ((e = launch_fft4_pass(false, 1 << logM1, channels, cols, s)) != hipSuccess) return e;
    if ((e = launch_fft4_pass(true, 1 << (logM - logM1), channels, rows, s)) != hipSuccess) return e;
    if (real && !inv) {
        hipLaunchKernelGGL(k_fft4_r2c_post, dim3(bins, (unsigned)channels), dim3(256), 0, s, Bs, M, a.out, a.stride,
                           logM, a.mode);
        return hipGetLastError();
    }
    return hipSuccess;
}

hipError_t launch_fft(int channels, const FftArgs &a, hipStream_t s)
{
    const bool real = a.mode >= 2;
    const int M = real ? (1 << (a.log2len - 1)) : (1 << a.log2len);
    if (M > kFftLdsMax) return launch_fft4(channels, a, s);
    switch (M) {
#define TANX_FFT_CASE(V) case V: hipLaunchKernelGGL(k_fft<V>, dim3(channels), dim3(fft_threads(V)), 0, s, a); break;
        TANX_FFT_CASE(1) TANX_FFT_CASE(2) TANX_FFT_CASE(4) TANX_FFT_CASE(8) TANX_FFT_CASE(16)
        TANX_FFT_CASE(32) TANX_FFT_CASE(64) TANX_FFT_CASE(128) TANX_FFT_CASE(256) TANX_FFT_CASE(512)
        TANX_FFT_CASE(1024) TANX_FFT_CASE(2048)
#undef TANX_FFT_CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_td(const TdArgs &a, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1)
{
    if (a.C == 0 || a.n == 0) return hipSuccess;
    if (a.len > kMaxTdLen || a.n > a.B) return hipErrorInvalidValue;
    const size_t smem = sizeof(float) * ((size_t)a.len - 1 + a.B + 2 * (size_t)a.len);
    hipExtLaunchKernelGGL(k_td, dim3(a.C), dim3(256), smem, s, ev0, ev1, 0, a);
    return hipGetLastError();
}

hipError_t launch_td_load(const float *ir, long long stride, long long n, const int *chan, int nchan, float *H,
                          int len, hipStream_t s)
{
    if (nchan == 0) return hipSuccess;
    const int g = (len + 255) / 256;
    const unsigned gx = (unsigned)(g > 64 ? 64 : g);
    hipLaunchKernelGGL(k_td_load, dim3(gx, nchan), dim3(256), 0, s, ir, stride, n, chan, H, len);
    return hipGetLastError();
}

hipError_t launch_dft_mfma(const float *x, long long in_stride, float *y, long long out_stride, int C, int n,
                           const void *fh, const void *fl, int split, hipStream_t s)
{
    if (C <= 0) return hipSuccess;
    if (n < 32 || (n & 31) || (split != 1 && split != 3)) return hipErrorInvalidValue;
    const int cols = n + 2;
    const unsigned blocks = (unsigned)(((C + 31) / 32) * ((cols + 31) / 32));
    hipLaunchKernelGGL(k_dft_mfma, dim3(blocks), dim3(64 * kDftWaves), 0, s, x, in_stride, y, out_stride, C, n, cols,
                       static_cast<const bf16x8_t *>(fh), static_cast<const bf16x8_t *>(fl), split);
    return hipGetLastError();
}

static unsigned io_blocks(long long n)
{
    long long b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    return (unsigned)(b < 1 ? 1 : b);
}

hipError_t launch_mix(const float *const *d_ins, const float *d_in, long long stride, int channels, long long n,
                      float *d_out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mix, dim3(io_blocks(n)), dim3(256), 0, s, d_ins, d_in, stride, channels, n, d_out);
    return hipGetLastError();
}

hipError_t launch_convert(const void *in, long long in_step, long long in_off, int in_type, void *out,
                          long long out_step, long long out_off, int out_type, long long n, float gain, int *d_flag,
                          hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const dim3 g(io_blocks(n)), t(256);
    if (in_type == 0 && out_type == 1)
        hipLaunchKernelGGL((k_convert<0, 1>), g, t, 0, s, in, in_step, in_off, out, out_step, out_off, n, gain, d_flag);
    else if (in_type == 1 && out_type == 0)
        hipLaunchKernelGGL((k_convert<1, 0>), g, t, 0, s, in, in_step, in_off, out, out_step, out_off, n, gain, d_flag);
    else if (in_type == 0 && out_type == 0)
        hipLaunchKernelGGL((k_convert<0, 0>), g, t, 0, s, in, in_step, in_off, out, out_step, out_off, n, gain, d_flag);
    else if (in_type == 1 && out_type == 1)
        hipLaunchKernelGGL((k_convert<1, 1>), g, t, 0, s, in, in_step, in_off, out, out_step, out_off, n, gain, d_flag);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_zero(void *p, size_t bytes, hipStream_t s)
{
    if (bytes == 0) return hipSuccess;
    if (bytes % 16 != 0 || (reinterpret_cast<uintptr_t>(p) & 15)) return hipErrorInvalidValue;
    const size_t n4 = bytes / 16;
    size_t blocks = (n4 + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, s, reinterpret_cast<float4 *>(p), n4);
    return hipGetLastError();
}

}  // namespace tanx
