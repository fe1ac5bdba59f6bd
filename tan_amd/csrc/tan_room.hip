// Room impulse responses by image sources on the GPU (TrueAudioVR's
// GenerateRoomResponse.cl and Fill.cl, driven as TrueAudioVRimpl::
// generateRoomResponse does, TrueAudioVR.cpp:659-778, 1038-1125).
//
// A shoebox room's response at one ear is the sum over mirror images of the
// source: image (ix, iy, iz) arrives at sample 1 + (int)(d / 340 * fs) with
// gain amplitude * dr, where amplitude is the product of the wall damping
// factors raised to the image's reflection counts and dr the 1/d distance
// law capped at max_gain.  Early images (arrival before hrtf_len) are
// shaped by the head filter lp + hf * hp, hf = (1 + cos(angle to the ear
// direction)) / 2; later ones add a single tap.  Contributions are summed
// in 2^26 fixed point with 32-bit integer atomics, as the reference does,
// so the sum does not depend on the order of the additions.
//
// Design for the MI355X: one grid row per ear (many rooms, sources and
// listeners in one launch).  Early images (the few near the ear) run one
// thread per image, and a wave adds their 64-tap head filters
// lane-parallel: each lane owns one tap, so one image's filter is one
// coalesced atomic instruction instead of 64 serial ones.  Late images (the
// vast majority, one tap each) are summed per window of output samples in
// LDS: the images of a window lie in a spherical shell, found row by row of
// the image lattice, so they need no global atomics at all.
//
// Arithmetic is IEEE single precision with no contraction, in the
// reference's operation order; integer powers of the damping factors are
// taken by binary powering (exactly reproducible on the host), where the
// reference calls native_powr.  Float-to-integer conversions saturate and
// map NaN to 0 (the reference's convert_int leaves those cases undefined).
#include "tan_internal.h"
#include "../../include/tan_mi355x.h"

#include <cmath>

// every expression in this file rounds each operation (device and host)
#pragma clang fp contract(off)

namespace tanx {
namespace {

__device__ __forceinline__ float ipow(float b, int e)
{
    float r = 1.0f;
    while (e > 0) {
        if (e & 1) r *= b;
        b *= b;
        e >>= 1;
    }
    return r;
}

__device__ __forceinline__ int f2i_sat(float v)
{
    if (!(v == v)) return 0;
    if (v >= 2147483647.0f) return 2147483647;
    if (v <= -2147483648.0f) return (-2147483647 - 1);
    return (int)v;
}

// Reflection counts of image index i on the two walls of its axis
// (GenerateRoomResponse.cl:238-244): first the wall at the axis' far side
// (right, top, front), then the near one (left, bottom, back).
__device__ __forceinline__ void bounces(int i, int &far_wall, int &near_wall)
{
    near_wall = i > 0 ? i / 2 : (1 - i) / 2;
    far_wall = i > 0 ? (i + 1) / 2 : (-i) / 2;
}

// Image coordinate on one axis: a * size + b * src, a = i (even) or i + 1
// (odd), b = +1 (even) or -1 (odd) (GenerateRoomResponse.cl:215-234).
__device__ __forceinline__ float image_pos(int i, float size, float src)
{
    const bool even = (i % 2) == 0;
    const float a = even ? (float)i : (float)(i + 1);
    const float b = even ? 1.0f : -1.0f;
    return a * size + b * src;
}

// One image's contribution (GenerateRoomResponse.cl:246-323): returns 0
// (none), 1 (one tap at *fi of (int)s) or 2 (taps head-filtered taps from
// *fi of (int)(s * (lp[i] + hf * hp[i]))).  Every caller evaluates an image
// with this same code, so its contribution is the same bits in every kernel.
template <class Params>
__device__ __forceinline__ int image_eval(const Params &E, int ix, int iy, int iz, int &fi, int &taps, float &s,
                                          float &hf)
{
    if (E.suppress_direct && ix == 0 && iy == 0 && iz == 0) return 0;
    const float rx = image_pos(ix, E.dim[0], E.src[0]);
    const float ry = image_pos(iy, E.dim[1], E.src[1]);
    const float rz = image_pos(iz, E.dim[2], E.src[2]);
    int rf, lf, tp, bt, fr, bk;
    bounces(ix, rf, lf);
    bounces(iy, tp, bt);
    bounces(iz, fr, bk);
    const float amp = ipow(E.damp[0], rf) * ipow(E.damp[1], lf) * ipow(E.damp[2], tp) * ipow(E.damp[3], bt) *
                      ipow(E.damp[4], fr) * ipow(E.damp[5], bk);
    const float dx = rx - E.head[0];
    const float dy = ry - E.head[1];
    const float dz = rz - E.head[2];
    const float d = sqrtf(dx * dx + dy * dy + dz * dz);
    const int q = f2i_sat((d / 340.0f) * (float)E.fs);  // arrival: sample 1 + q
    if (q < 0 || q >= E.len - 1) return 0;
    fi = q + 1;
    const float dr = (d <= E.d_min) ? E.max_gain : E.max_gain * E.d_min / d;
    const int room_left = E.len - fi;
    taps = E.filt_len > room_left ? room_left : E.filt_len;
    const float dp = E.ev[0] * dx + E.ev[1] * dy + E.ev[2] * dz;
    const float cosa = dp / (E.ev_norm * d);
    hf = (cosa + 1.0f) / 2.0f;
    s = kRoomFixedOne * amp * dr;
    return fi < E.hrtf_len ? 2 : 1;
}

// Early images: one thread per image of the early box (every image that can
// arrive before hrtf_len lies in it); the wave adds each early image's head
// filter lane-parallel, one lane per tap.  Late images are left to the
// window kernel.
__global__ void __launch_bounds__(256) k_room_early(const RoomEar *__restrict__ ears, unsigned *__restrict__ acc,
                                                    int len)
{
    const RoomEar &E = ears[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int wx = E.hi[0] - E.lo[0] + 1, wy = E.hi[1] - E.lo[1] + 1, wz = E.hi[2] - E.lo[2] + 1;
    const long long total = (wx > 0 && wy > 0 && wz > 0) ? (long long)wx * wy * wz : 0;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned *row = acc + (size_t)blockIdx.y * (size_t)len;

    bool early = false;
    int idx = 0, taps = 0;
    float s = 0.0f, hf = 0.0f;
    if (id < total) {
        const int ix = (int)(id % wx) + E.lo[0];
        const long long t = id / wx;
        const int iy = (int)(t % wy) + E.lo[1];
        const int iz = (int)(t / wy) + E.lo[2];
        early = image_eval(E, ix, iy, iz, idx, taps, s, hf) == 2 && taps > 0;
    }
    const float lp = lane < E.filt_len ? E.lp[lane] : 0.0f;
    const float hp = lane < E.filt_len ? E.hp[lane] : 0.0f;
    unsigned long long pend = __ballot(early);
    while (pend) {
        const int src = __ffsll((long long)pend) - 1;
        pend &= pend - 1;
        const int j = __shfl(idx, src);
        const int nt = __shfl(taps, src);
        const float sj = __shfl(s, src);
        const float hj = __shfl(hf, src);
        if (lane < nt) atomicAdd(row + j + lane, (unsigned)f2i_sat(sj * (lp + hj * hp)));
    }
}

// The per-ear scalars an image needs, held in registers by the window
// kernel (its loops would otherwise reload them around every LDS atomic).
struct ImageParams {
    float src[3], head[3], ev[3], ev_norm, dim[3], damp[6], max_gain, d_min;
    int fs, len, hrtf_len, filt_len, suppress_direct;
};

__device__ __forceinline__ ImageParams image_params(const RoomEar &E)
{
    ImageParams p;
    for (int k = 0; k < 3; k++) {
        p.src[k] = E.src[k];
        p.head[k] = E.head[k];
        p.ev[k] = E.ev[k];
        p.dim[k] = E.dim[k];
    }
    for (int k = 0; k < 6; k++) p.damp[k] = E.damp[k];
    p.ev_norm = E.ev_norm;
    p.max_gain = E.max_gain;
    p.d_min = E.d_min;
    p.fs = E.fs;
    p.len = E.len;
    p.hrtf_len = E.hrtf_len;
    p.filt_len = E.filt_len;
    p.suppress_direct = E.suppress_direct;
    return p;
}

// Late images of one window of kRoomWindow output samples [w0, w0 +
// kRoomWindow), w0 = wstart + window * kRoomWindow, and one slice of the
// image lattice's rows (iy, iz): the images arriving in the window lie in a
// spherical shell around the ear.  A thread takes a row, finds the cells ix
// whose images can lie in the shell (its x extent at that row; the radii
// carry a margin far above the rounding of the arrival rule, and every
// candidate is then evaluated exactly), and adds the single-tap
// contributions arriving in the window to an LDS sum.  The workgroup then
// adds its nonzero sums to acc: one atomic per bin instead of one per image.
constexpr int kPowTable = 512;  // damping powers 0..511 per wall in LDS

__global__ void __launch_bounds__(256) k_room_late(const RoomEar *__restrict__ ears, unsigned *__restrict__ acc,
                                                   int len, int slices)
{
    __shared__ unsigned sum[kRoomWindow];
    __shared__ float pw[6][kPowTable];  // pw[k][e] = ipow(damp[k], e): the same bits as ipow
    const RoomEar &E = ears[blockIdx.y];
    const int window = (int)blockIdx.x / slices, slice = (int)blockIdx.x % slices;
    const int w0 = E.wstart + window * kRoomWindow;
    if (w0 >= len) return;  // uniform over the workgroup
    const int w1 = min(w0 + kRoomWindow, len);
    const ImageParams P = image_params(E);
    const int n0 = E.n[0], n1 = E.n[1], n2 = E.n[2];
    // reflection counts are at most n per axis
    const bool tables = max(n0, max(n1, n2)) < kPowTable;
    for (int i = threadIdx.x; i < kRoomWindow; i += blockDim.x) sum[i] = 0u;
    if (tables) {
        const int ne = max(n0, max(n1, n2)) + 1;
        for (int i = threadIdx.x; i < 6 * ne; i += blockDim.x) pw[i / ne][i % ne] = ipow(P.damp[i / ne], i % ne);
    }
    __syncthreads();
    auto power = [&](int k, int e) { return tables ? pw[k][e] : ipow(P.damp[k], e); };

    const float m = 0.05f;
    const float dlo = fmaxf(0.0f, (float)(w0 - 1) * 340.0f / (float)P.fs - m);
    const float dhi = (float)w1 * 340.0f / (float)P.fs + m;
    const int wy = 2 * n1 + 1, rows = wy * (2 * n2 + 1);
    const float W = P.dim[0], hx = P.head[0];
    for (int r = slice + slices * (int)threadIdx.x; r < rows; r += slices * (int)blockDim.x) {
        const int iy = r % wy - n1;
        const int iz = r / wy - n2;
        const float dy = image_pos(iy, P.dim[1], P.src[1]) - P.head[1];
        const float dz = image_pos(iz, P.dim[2], P.src[2]) - P.head[2];
        const float D2 = dy * dy + dz * dz;
        if (D2 > dhi * dhi) continue;
        const float xh = sqrtf(fmaxf(0.0f, dhi * dhi - D2)) + m;
        const float xl = fmaxf(0.0f, sqrtf(fmaxf(0.0f, dlo * dlo - D2)) - m);
        // cells [ix W, (ix + 1) W] meeting [hx - xh, hx - xl] or [hx + xl, hx + xh]
        int a0 = (int)floorf((hx - xh) / W), a1 = (int)floorf((hx - xl) / W);
        int b0 = (int)floorf((hx + xl) / W), b1 = (int)floorf((hx + xh) / W);
        if (a1 >= b0 - 1) {
            a1 = b1;
            b0 = 1;
            b1 = 0;  // one merged interval
        }
        a0 = max(a0, -n0);
        a1 = min(a1, n0);
        b0 = max(b0, -n0);
        b1 = min(b1, n0);
        // the row's factors (image_eval's arithmetic, hoisted out of the ix loop)
        int tp, bt, fr, bk;
        bounces(iy, tp, bt);
        bounces(iz, fr, bk);
        const float pT = power(2, tp), pB = power(3, bt), pF = power(4, fr), pK = power(5, bk);
        for (int pass = 0; pass < 2; pass++) {
            const int lo = pass ? b0 : a0, hi = pass ? b1 : a1;
            for (int ix = lo; ix <= hi; ix++) {
                if (P.suppress_direct && ix == 0 && iy == 0 && iz == 0) continue;
                int rf, lf;
                bounces(ix, rf, lf);
                const float amp = power(0, rf) * power(1, lf) * pT * pB * pF * pK;
                const float dx = image_pos(ix, W, P.src[0]) - hx;
                const float d = sqrtf(dx * dx + dy * dy + dz * dz);
                const int q = f2i_sat((d / 340.0f) * (float)P.fs);
                // late (one tap): arrival 1 + q in [hrtf_len, len), and in the window
                if (q < 0 || q >= P.len - 1) continue;
                const int fi = q + 1;
                if (fi < P.hrtf_len || fi < w0 || fi >= w1) continue;
                const float dr = (d <= P.d_min) ? P.max_gain : P.max_gain * P.d_min / d;
                atomicAdd(&sum[fi - w0], (unsigned)f2i_sat(kRoomFixedOne * amp * dr));
            }
        }
    }
    __syncthreads();
    unsigned *arow = acc + (size_t)blockIdx.y * (size_t)len;
    for (int f = w0 + threadIdx.x; f < w1; f += blockDim.x) {
        const unsigned v = sum[f - w0];
        if (v) atomicAdd(arow + f, v);
    }
}

// Fill.cl: the fixed-point sums to float (/ 2^26, read as signed), and the
// accumulator cleared for the next call.
__global__ void __launch_bounds__(256) k_room_fill(unsigned *__restrict__ acc, int len, float *__restrict__ out,
                                                   long long out_stride)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    unsigned *row = acc + (size_t)blockIdx.y * (size_t)len;
    out[(size_t)blockIdx.y * (size_t)out_stride + i] = (float)(int)row[i] / kRoomFixedOne;
    row[i] = 0u;
}

}  // namespace

// Kernel arguments of both ears of one response: TrueAudioVRimpl::
// generateRoomResponse's host part (TrueAudioVR.cpp:659-778) with rotMtx
// (TrueAudioVR.cpp:614-651).  Returns the image count per ear in *images.
int32_t room_ears(const tanx_room *room, const tanx_source *source, const tanx_listener *ears, int fs, int len,
                  uint32_t flags, int max_bounces, RoomEar out[2], long long *images)
{
    const float S = 340.0f;  // AmdTrueAudioVR::S
    if (!(room->width > 0.0f) || !(room->height > 0.0f) || !(room->length > 0.0f) || fs <= 0 || len <= 0)
        return TANX_INVALID_ARG;
    const float damps[6] = {room->damp_right, room->damp_left,  room->damp_top,
                            room->damp_bottom, room->damp_front, room->damp_back};
    for (float d : damps)
        if (!(d >= 0.0f && d <= 1.0f)) return TANX_INVALID_ARG;
    if (ears->filter_length < 0 || ears->filter_length > kHeadFilterMax) return TANX_INVALID_ARG;
    if ((flags & TANX_ROOM_LIMIT_BOUNCES) && max_bounces < 0) return TANX_INVALID_ARG;

    float sx = source->x, sy = source->y, sz = source->z;
    if (sx > room->width) sx = room->width;
    if (sx < 0) sx = 0;
    if (sy > room->height) sy = room->height;
    if (sy < 0) sy = 0;
    if (sz > room->length) sz = room->length;
    if (sz < 0) sz = 0;

    // ear offsets and directions, rotated by yaw, pitch, roll (degrees)
    float e[4][3] = {{ears->ear_spacing / 2, 0.0f, 0.0f},
                     {-ears->ear_spacing / 2, 0.0f, 0.0f},
                     {0.0f, 0.0f, 0.0f},
                     {0.0f, 0.0f, 0.0f}};
    e[2][0] = e[0][0];
    e[2][2] = -e[0][0];
    e[3][0] = e[1][0];
    e[3][2] = e[1][0];
    const float pi = (float)3.141592653589793;
    const float sinY = std::sin(ears->yaw * pi / 180), cosY = std::cos(ears->yaw * pi / 180);
    const float sinP = std::sin(ears->pitch * pi / 180), cosP = std::cos(ears->pitch * pi / 180);
    const float sinR = std::sin(ears->roll * pi / 180), cosR = std::cos(ears->roll * pi / 180);
    const float m[3][3] = {{cosR * cosY - sinR * sinP * sinY, -sinR * cosP, cosR * sinY + sinR * sinP * cosY},
                           {sinR * cosY + cosR * sinP * sinY, cosR * cosP, sinR * sinY - cosR * sinP * cosY},
                           {-cosP * sinY, sinP, cosP * cosY}};
    for (auto &v : e) {
        const float x = v[0], y = v[1], z = v[2];
        v[0] = x * m[0][0] + y * m[0][1] + z * m[0][2];
        v[1] = x * m[1][0] + y * m[1][1] + z * m[1][2];
        v[2] = x * m[2][0] + y * m[2][1] + z * m[2][2];
    }

    const int nsw = (int)((room->width / S) * (float)fs);
    const int nsh = (int)((room->height / S) * (float)fs);
    const int nsl = (int)((room->length / S) * (float)fs);
    if (nsw < 1 || nsh < 1 || nsl < 1) return TANX_INVALID_ARG;  // the reference divides by these
    int n[3] = {1 + len / nsw, 1 + len / nsh, 1 + len / nsl};
    if (flags & TANX_ROOM_LIMIT_BOUNCES)
        for (int &v : n) v = v > max_bounces ? max_bounces : v;
    if ((double)(2 * n[0] + 1) * (double)(2 * n[1] + 1) * (double)(2 * n[2] + 1) > 2147483647.0)
        return TANX_OUT_OF_RANGE;
    *images = (long long)(2 * n[0] + 1) * (2 * n[1] + 1) * (2 * n[2] + 1);

    for (int ch = 0; ch < 2; ch++) {
        RoomEar &E = out[ch];
        E.src[0] = sx;
        E.src[1] = sy;
        E.src[2] = sz;
        for (int k = 0; k < 3; k++) {
            E.dim[k] = k == 0 ? room->width : k == 1 ? room->height : room->length;
            E.n[k] = n[k];
        }
        E.head[0] = ears->head_x + e[ch][0];
        E.head[1] = ears->head_y + e[ch][1];
        E.head[2] = ears->head_z + e[ch][2];
        for (int k = 0; k < 3; k++) E.ev[k] = e[2 + ch][k];
        E.ev_norm = std::sqrt(E.ev[0] * E.ev[0] + E.ev[1] * E.ev[1] + E.ev[2] * E.ev[2]);
        for (int k = 0; k < 6; k++) E.damp[k] = damps[k];
        E.max_gain = 2.0f;
        E.d_min = 2 * ears->ear_spacing;
        E.fs = fs;
        E.len = len;
        E.hrtf_len = 2 * (nsw + nsh + nsl);  // the reference's cutoff for head filtering of echoes
        E.filt_len = ears->filter_length;
        E.suppress_direct = (flags & TANX_ROOM_SUPPRESS_DIRECT) ? 1 : 0;
        for (int k = 0; k < kHeadFilterMax; k++) {
            E.lp[k] = k < ears->filter_length ? ears->low_pass[k] : 0.0f;
            E.hp[k] = k < ears->filter_length ? ears->high_pass[k] : 0.0f;
        }
        // Images arriving before min(hrtf_len, len) lie within R of the ear
        // (1 m above the arrival rule's radius); image ix of an axis lies in
        // cell [ix size, (ix + 1) size], so the box below holds them all.
        E.wstart = E.hrtf_len < len ? E.hrtf_len : len;
        const double R = (double)E.wstart * 340.0 / (double)fs + 1.0;
        for (int k = 0; k < 3; k++) {
            const double size = E.dim[k], h = E.head[k];
            const double lo = std::floor((h - R) / size) - 1.0, hi = std::floor((h + R) / size) + 1.0;
            E.lo[k] = lo < -n[k] ? -n[k] : (int)lo;
            E.hi[k] = hi > n[k] ? n[k] : (int)hi;
        }
    }
    return TANX_OK;
}

hipError_t launch_room(const RoomEar *d_ears, int nears, long long max_early, int max_windows, int len,
                       unsigned *acc, float *out, long long out_stride, hipStream_t s)
{
    if (nears <= 0 || len <= 0) return hipSuccess;
    const long long blocks = (max_early + 255) / 256;
    if (blocks > 0x7fffffffLL || nears > 65535 || max_windows < 0) return hipErrorInvalidValue;
    if (blocks > 0) {
        hipLaunchKernelGGL(k_room_early, dim3((unsigned)blocks, (unsigned)nears), dim3(256), 0, s, d_ears, acc, len);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (max_windows > 0) {
        // row slices per window: about 2,048 workgroups in all, at most 32 slices
        long long per = 2048LL / ((long long)max_windows * nears);
        const int slices = (int)(per < 1 ? 1 : per > 32 ? 32 : per);
        hipLaunchKernelGGL(k_room_late, dim3((unsigned)(max_windows * slices), (unsigned)nears), dim3(256), 0, s,
                           d_ears, acc, len, slices);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_room_fill, dim3((unsigned)((len + 255) / 256), (unsigned)nears), dim3(256), 0, s, acc, len,
                       out, out_stride);
    return hipGetLastError();
}

}  // namespace tanx

// generateSimpleHeadRelatedTransform (TrueAudioVR.cpp:361-415): a 64-tap
// low-pass for the head's acoustic shadow, corner 343 / (1.1 * spacing) Hz,
// and its complement (the two sum to an impulse).  The reference filters a
// 64-point complex spectrum of ones by d_j = fc / (fc + f_j), f_j =
// floor(j / 2) * 48000 / 64, and keeps the real part of the inverse FFT
// (scaled 1/64); here that real part is summed directly in FP64:
// lp[k] = (1/64) sum_j d_j cos(2 pi j k / 64).
extern "C" int32_t tanx_room_head_model(float ear_spacing, tanx_listener *listener)
{
    if (!listener) return TANX_INVALID_POINTER;
    if (!(ear_spacing > 0.0f)) return TANX_INVALID_ARG;
    constexpr int n = tanx::kHeadFilterMax;
    const float fc = (float)(343.0 / (1.10 * (double)ear_spacing));
    const float in_delta = (float)(1.0 / (float)48000);
    float dj[n];
    for (int j = 0; j < n; j++) {
        const float f = (float)((float)(j / 2) / ((float)n * in_delta));
        dj[j] = (float)(fc / (fc + f));
    }
    listener->filter_length = n;
    for (int k = 0; k < n; k++) {
        double acc = 0.0;
        for (int j = 0; j < n; j++) acc += (double)dj[j] * std::cos(2.0 * M_PI * (double)((j * k) % n) / (double)n);
        listener->low_pass[k] = (float)(acc / n);
    }
    listener->high_pass[0] = 1.0f;
    for (int k = 1; k < n; k++) listener->high_pass[k] = 0.0f;
    for (int k = 0; k < n; k++) listener->high_pass[k] -= listener->low_pass[k];
    return TANX_OK;
}
