/*
 * tan_oracle.c -- CPU oracle for the TAN time-varying convolution path.
 *
 * TEST INFRASTRUCTURE ONLY (checker and CPU baseline); never part of the
 * product library.  Reference citations are relative to the TAN checkout.
 *
 * Pinning of each part (what a plausible mistake would fail):
 *   orc_dft_f64, orc_direct_conv_f64 -- mathematical definitions; checked in
 *       tests against numpy float64 FFT/convolve and closed forms.
 *   orc_fft_transform -- checked against orc_dft_f64 on random inputs, impulse
 *       (flat spectrum), single cosine (N/2 at +-k), Parseval, round trip.
 *       FFTW's exact rounding: parity unpinned (FFTW absent).
 *   orc_conv_* -- checked against orc_direct_conv_f64 (steady state and
 *       intended crossfade), impulse response reproduction, linearity,
 *       channel independence, zero IR -> zero output; the planted mistakes
 *       of orc_set_mutation() must each fail one of those checks.
 */
#include "tan_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <pthread.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define AMF_OK 0
#define AMF_INVALID_ARG 4
#define AMF_NOT_FOUND 11
#define AMF_NOT_INITIALIZED 13
#define AMF_WRONG_STATE 15

#define FLAG_STOP_INPUT 0x01u
#define FLAG_FLUSH_STREAM 0x02u
#define OPFLAG_BLOCK_UNTIL_READY 0x01u

/* PARTITION_PAD_FFTREAL_PLANAR (ConvolutionImpl.h:50) */
#define PLANAR_PAD 16
/* planar imaginary plane offset in complex numbers: N/2 + 8
 * (FFTImpl.cpp:961-978, 1028-1043) */
#define PLANAR_IM_OFF(M) ((M) + 8)

static int g_mutation = 0;
void orc_set_mutation(int m) { g_mutation = m; }

int orc_num_procs(void)
{
#ifdef _OPENMP
    return omp_get_num_procs();
#else
    return 1;
#endif
}

int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n < 1) n = 1;
    omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ======================================================================= */
/* 1. FP64 references                                                       */
/* ======================================================================= */

void orc_dft_f64(int n_len, const double *in, double *out, int inverse)
{
    const double sgn = inverse ? 1.0 : -1.0;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n_len; k++) {
        double sr = 0.0, si = 0.0;
        for (int n = 0; n < n_len; n++) {
            /* exact integer reduction of n*k mod N keeps the angle small */
            long long nk = ((long long)n * k) % n_len;
            double ang = sgn * 2.0 * M_PI * (double)nk / (double)n_len;
            double c = cos(ang), s = sin(ang);
            double xr = in[2 * n], xi = in[2 * n + 1];
            sr += xr * c - xi * s;
            si += xr * s + xi * c;
        }
        out[2 * k] = sr;
        out[2 * k + 1] = si;
    }
}

void orc_direct_conv_f64(const float *x, long nx, const float *h, long nh,
                         double *y, long ny)
{
    #pragma omp parallel for schedule(static)
    for (long n = 0; n < ny; n++) {
        double acc = 0.0;
        long k0 = n - (nx - 1);
        if (k0 < 0) k0 = 0;
        long k1 = n < nh - 1 ? n : nh - 1;
        for (long k = k0; k <= k1; k++) acc += (double)h[k] * (double)x[n - k];
        y[n] = acc;
    }
}

/* ======================================================================= */
/* 2. FP32 FFT (own radix-2; twiddles computed in FP64, stored FP32)        */
/* ======================================================================= */

#define MAX_LOG2 21
static float *g_tw[MAX_LOG2 + 1]; /* g_tw[m]: exp(-2 pi i k / 2^m), k < 2^(m-1)+1 */
static pthread_mutex_t g_tw_lock = PTHREAD_MUTEX_INITIALIZER;

static const float *twiddles(int log2n)
{
    float *t = __atomic_load_n(&g_tw[log2n], __ATOMIC_ACQUIRE);
    if (t) return t;
    pthread_mutex_lock(&g_tw_lock);
    t = g_tw[log2n];
    if (!t) {
        long n = 1L << log2n;
        long cnt = n / 2 + 1;
        t = (float *)malloc(sizeof(float) * 2 * cnt);
        for (long k = 0; k < cnt; k++) {
            /* angle from an exact rational: k/n in [0, 1/2] */
            double ang = -2.0 * M_PI * (double)k / (double)n;
            t[2 * k] = (float)cos(ang);
            t[2 * k + 1] = (float)sin(ang);
        }
        __atomic_store_n(&g_tw[log2n], t, __ATOMIC_RELEASE);
    }
    pthread_mutex_unlock(&g_tw_lock);
    return t;
}

/* in-place unscaled complex FFT of 2^log2m points, interleaved. */
static void cfft(float *z, int log2m, int inverse)
{
    const long m = 1L << log2m;
    if (m == 1) return;
    const float *tw = twiddles(log2m);
    /* bit reversal */
    for (long i = 1, j = 0; i < m; i++) {
        long bit = m >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            float tr = z[2 * i], ti = z[2 * i + 1];
            z[2 * i] = z[2 * j];
            z[2 * i + 1] = z[2 * j + 1];
            z[2 * j] = tr;
            z[2 * j + 1] = ti;
        }
    }
    const float isgn = inverse ? -1.0f : 1.0f;
    for (long len = 2; len <= m; len <<= 1) {
        long half = len >> 1, step = m / len;
        for (long base = 0; base < m; base += len) {
            for (long j = 0; j < half; j++) {
                float wr = tw[2 * (j * step)], wi = isgn * tw[2 * (j * step) + 1];
                float *a = z + 2 * (base + j), *b = z + 2 * (base + j + half);
                float br = b[0] * wr - b[1] * wi;
                float bi = b[0] * wi + b[1] * wr;
                b[0] = a[0] - br;
                b[1] = a[1] - bi;
                a[0] += br;
                a[1] += bi;
            }
        }
    }
}

/* Real forward FFT of N = 2^log2n reals -> X[0..N/2] (re[k], im[k] arrays). */
static void rfft_fwd(const float *x, int log2n, float *re, float *im, float *tmp)
{
    const long n = 1L << log2n, m = n / 2;
    if (n == 1) { re[0] = x[0]; im[0] = 0.0f; return; }
    memcpy(tmp, x, sizeof(float) * n); /* z[j] = x[2j] + i x[2j+1] */
    cfft(tmp, log2n - 1, 0);
    const float *tw = twiddles(log2n);
    for (long k = 0; k <= m; k++) {
        long k1 = k % m, k2 = (m - k) % m;
        float zr = tmp[2 * k1], zi = tmp[2 * k1 + 1];
        float cr = tmp[2 * k2], ci = -tmp[2 * k2 + 1]; /* conj(Z[m-k]) */
        float er = 0.5f * (zr + cr), ei = 0.5f * (zi + ci);
        /* O = (Z - conj)/(2i) = (-i/2)(Z - conj) */
        float dr = zr - cr, di = zi - ci;
        float orr = 0.5f * di, oi = -0.5f * dr;
        float wr = tw[2 * k], wi = tw[2 * k + 1];
        re[k] = er + (wr * orr - wi * oi);
        im[k] = ei + (wr * oi + wi * orr);
    }
}

/* Real backward FFT: X[0..N/2] -> N reals scaled by 1/N.  Imag parts of the
 * DC and Nyquist bins are ignored, as a c2r transform does. */
static void rfft_bwd(const float *re, const float *im, int log2n, float *x, float *tmp)
{
    const long n = 1L << log2n, m = n / 2;
    if (n == 1) { x[0] = re[0]; return; }
    const float *tw = twiddles(log2n);
    for (long k = 0; k < m; k++) {
        float xr = re[k], xi = (k == 0) ? 0.0f : im[k];
        float yr = re[m - k], yi = (m - k == m || m - k == 0) ? 0.0f : im[m - k];
        yi = -yi; /* conj(X[m-k]) */
        float er = xr + yr, ei = xi + yi;
        float dr = xr - yr, di = xi - yi;
        float wr = tw[2 * k], wi = -tw[2 * k + 1]; /* conj(W^k) */
        float orr = dr * wr - di * wi, oi = dr * wi + di * wr;
        tmp[2 * k] = er - oi;     /* E + i O */
        tmp[2 * k + 1] = ei + orr;
    }
    cfft(tmp, log2n - 1, 1);
    const float s = (g_mutation == 4) ? 1.0f : 1.0f / (float)n;
    for (long j = 0; j < n; j++) x[j] = tmp[j] * s;
}

int orc_fft_transform(int direction, int log2len, int channels,
                      float *const *in, float *const *out)
{
    if (!in || !out || channels <= 0 || log2len <= 0 || log2len > MAX_LOG2 - 1)
        return AMF_INVALID_ARG;
    if (direction < 0 || direction > 5) return AMF_INVALID_ARG;
    /* the planar layout puts im at N/2+8 and ends at N+9 floats; a channel
     * buffer holds 2N floats, so planar needs N >= 16 */
    if (direction >= ORC_FFT_R2C_PLANAR_FORWARD && log2len < 4) return AMF_INVALID_ARG;
    const long n = 1L << log2len, m = n / 2;
    int err = 0;
    #pragma omp parallel for schedule(static)
    for (int c = 0; c < channels; c++) {
        float *tmp = (float *)malloc(sizeof(float) * (2 * n + 32));
        float *re = (float *)malloc(sizeof(float) * (m + 2));
        float *im = (float *)malloc(sizeof(float) * (m + 2));
        float *src = in[c], *dst = out[c];
        switch (direction) {
        case ORC_FFT_FORWARD:
        case ORC_FFT_BACKWARD: {
            memcpy(tmp, src, sizeof(float) * 2 * n);
            cfft(tmp, log2len, direction == ORC_FFT_BACKWARD);
            float s = direction == ORC_FFT_BACKWARD ? 1.0f / (float)n : 1.0f;
            for (long j = 0; j < 2 * n; j++) dst[j] = tmp[j] * s;
            break;
        }
        case ORC_FFT_R2C_FORWARD:
        case ORC_FFT_R2C_PLANAR_FORWARD: {
            rfft_fwd(src, log2len, re, im, tmp);
            if (direction == ORC_FFT_R2C_FORWARD) {
                for (long k = 0; k <= m; k++) { dst[2 * k] = re[k]; dst[2 * k + 1] = im[k]; }
            } else {
                float *pim = dst + PLANAR_IM_OFF(m);
                for (long k = 0; k <= m; k++) { dst[k] = re[k]; pim[k] = im[k]; }
            }
            break;
        }
        case ORC_FFT_C2R_BACKWARD:
        case ORC_FFT_C2R_PLANAR_BACKWARD: {
            if (direction == ORC_FFT_C2R_BACKWARD) {
                for (long k = 0; k <= m; k++) { re[k] = src[2 * k]; im[k] = src[2 * k + 1]; }
            } else {
                const float *pim = src + PLANAR_IM_OFF(m);
                for (long k = 0; k <= m; k++) { re[k] = src[k]; im[k] = pim[k]; }
            }
            float *xo = (float *)malloc(sizeof(float) * n);
            rfft_bwd(re, im, log2len, xo, tmp);
            memcpy(dst, xo, sizeof(float) * n);
            free(xo);
            break;
        }
        default:
            err = 1;
        }
        free(tmp);
        free(re);
        free(im);
    }
    return err ? AMF_INVALID_ARG : AMF_OK;
}

/* ======================================================================= */
/* 3. Uniform-partitioned convolution (UPOLS with FDL), restated            */
/* ======================================================================= */

struct orc_conv {
    int L, B, C;
    int len;     /* m_length: next power of two >= L (ConvolutionImpl.cpp:1230-1235) */
    int P;       /* partitions: m_length / B */
    int log2N;   /* log2 of the 2B-point FFT */
    int partLen; /* 2B + PLANAR_PAD floats per partition */
    long bufLen; /* 2 * (2 * m_length + PAD * P) (ConvolutionImpl.cpp:1598) */
    int xmode, fmode;

    float **staged;   /* [C][len] time-domain IR staged by UpdateResponseTD */
    int *stagedSet;   /* channel updated in the pending update */
    float **filt[3];  /* [3][C][bufLen] partitioned planar IR spectra */
    float **data;     /* [C][bufLen] FDL ring of input spectra */
    float **overlap;  /* [C][2B] */
    float **ovNew;    /* [C][B] intended crossfade: new-filter overlap */
    float **outS;     /* [C][2B+PAD]: tail accumulator for the next head */
    float **outX;     /* [C][2B+PAD]: second tail (crossfade) */
    float **acc;      /* [C][2B+PAD] work */
    float **xfade;    /* [C][B] new-filter output in the crossfade block */
    float **work;     /* [C][4B+64] FFT scratch */

    int curPart;                 /* m_currentDataPartition */
    int idxFilter, idxPrev, idxUpdate;
    int pending, readyEvent, delayed, crossFading, lastXfade;

    unsigned char *available, *flushed;
    int *tailLeft;
};

static float **alloc_rows(int rows, long cols)
{
    float **p = (float **)calloc((size_t)rows, sizeof(float *));
    for (int i = 0; i < rows; i++) p[i] = (float *)calloc((size_t)cols, sizeof(float));
    return p;
}

static void free_rows(float **p, int rows)
{
    if (!p) return;
    for (int i = 0; i < rows; i++) free(p[i]);
    free(p);
}

orc_conv *orc_conv_create(int L, int B, int C, int xfade_mode, int finalize_mode)
{
    if (L <= 0 || B <= 0 || C <= 0 || (B & (B - 1))) return NULL;
    orc_conv *c = (orc_conv *)calloc(1, sizeof(orc_conv));
    c->L = L; c->B = B; c->C = C;
    c->xmode = xfade_mode; c->fmode = finalize_mode;
    int len = 1;
    while (len < L) len <<= 1;
    if (len < B) len = B;
    c->len = len;
    c->P = len / B;
    int l2 = 0;
    while ((1 << l2) < 2 * B) l2++;
    c->log2N = l2;
    c->partLen = 2 * B + PLANAR_PAD;
    c->bufLen = (long)c->P * c->partLen + 64;
    c->staged = alloc_rows(C, len);
    c->stagedSet = (int *)calloc(C, sizeof(int));
    for (int s = 0; s < 3; s++) c->filt[s] = alloc_rows(C, c->bufLen);
    c->data = alloc_rows(C, c->bufLen);
    c->overlap = alloc_rows(C, 2 * B);
    c->ovNew = alloc_rows(C, B);
    c->outS = alloc_rows(C, c->partLen);
    c->outX = alloc_rows(C, c->partLen);
    c->acc = alloc_rows(C, c->partLen);
    c->xfade = alloc_rows(C, B);
    c->work = alloc_rows(C, 4 * B + 64);
    /* initial slots (ConvolutionImpl.cpp:115-117): current 1, previous 0,
     * update 2; slot 1 holds zeros, so output is zero until the first fade */
    c->idxFilter = 1; c->idxPrev = 0; c->idxUpdate = 2;
    c->available = (unsigned char *)calloc(C, 1);
    c->flushed = (unsigned char *)calloc(C, 1);
    c->tailLeft = (int *)calloc(C, sizeof(int));
    return c;
}

void orc_conv_destroy(orc_conv *c)
{
    if (!c) return;
    int C = c->C;
    free_rows(c->staged, C);
    free(c->stagedSet);
    for (int s = 0; s < 3; s++) free_rows(c->filt[s], C);
    free_rows(c->data, C);
    free_rows(c->overlap, C);
    free_rows(c->ovNew, C);
    free_rows(c->outS, C);
    free_rows(c->outX, C);
    free_rows(c->acc, C);
    free_rows(c->xfade, C);
    free_rows(c->work, C);
    free(c->available);
    free(c->flushed);
    free(c->tailLeft);
    free(c);
}

/* planar R2C of one 2B-point partition in place (FFTImpl.cpp:1028-1043) */
static void part_r2c(orc_conv *c, float *part, float *work)
{
    const int B = c->B;
    float *re = work, *im = work + (B + 8), *tmp = work + 2 * B + 32;
    rfft_fwd(part, c->log2N, re, im, tmp);
    memcpy(part, re, sizeof(float) * (B + 1));
    memcpy(part + PLANAR_IM_OFF(B), im, sizeof(float) * (B + 1));
}

/* planar C2R in place: part -> 2B reals scaled 1/N (FFTImpl.cpp:1028-1051) */
static void part_c2r(orc_conv *c, float *part, float *work)
{
    const int B = c->B;
    float *xo = work, *tmp = work + 2 * B + 16;
    rfft_bwd(part, part + PLANAR_IM_OFF(B), c->log2N, xo, tmp);
    memcpy(part, xo, sizeof(float) * 2 * B);
}

/* PlanarComplexMultiplyAccumulate, one channel, the AVX2 path's arithmetic
 * (MathImpl.cpp:495-530): cr += fmsub(ar, br, ai*bi); ci += fmadd(ar, bi, ai*br).
 * count = B+8, riPlaneSpacing = B+8 (ConvolutionImpl.cpp:2496, 2622). */
static void planar_cmac(const float *a, const float *b, float *acc, int count, int sp)
{
    const float *ar = a, *ai = a + sp, *br = b, *bi = b + sp;
    float *cr = acc, *ci = acc + sp;
    const float conj = (g_mutation == 1) ? -1.0f : 1.0f;
    for (int k = 0; k < count; k++) {
        float bik = conj * bi[k];
        cr[k] = fmaf(ar[k], br[k], -(ai[k] * bik)) + cr[k];
        ci[k] = fmaf(ar[k], bik, ai[k] * br[k]) + ci[k];
    }
}

static inline float *part_ptr(float *row, const orc_conv *c, int i)
{
    return row + (long)i * c->partLen;
}

/* Flush (ConvolutionImpl.cpp:1261-1325): zero the channel's overlap and FDL.
 * The pending tail is zeroed too, so the flushed block starts from silence. */
static void flush_channel(orc_conv *c, int ch)
{
    memset(c->overlap[ch], 0, sizeof(float) * 2 * c->B);
    memset(c->data[ch], 0, sizeof(float) * c->bufLen);
    memset(c->outS[ch], 0, sizeof(float) * c->partLen);
    memset(c->outX[ch], 0, sizeof(float) * c->partLen);
    memset(c->ovNew[ch], 0, sizeof(float) * c->B);
}

/* Expand a staged time-domain IR into P planar spectra (update thread,
 * ConvolutionImpl.cpp:2080-2147): partition i = [h_i (B), 0 (B + pad)]. */
static void transform_ir(orc_conv *c, int slot, int ch)
{
    const int B = c->B;
    float *dst = c->filt[slot][ch];
    memset(dst, 0, sizeof(float) * c->bufLen);
    for (int i = 0; i < c->P; i++) {
        float *p = part_ptr(dst, c, i);
        memcpy(p, c->staged[ch] + (long)i * B, sizeof(float) * B);
        part_r2c(c, p, c->work[ch]);
    }
}

/* ovlNUPProcessCPU with M = 1 (ConvolutionImpl.cpp:2380-2535).  Writes the
 * block output into out[ch]. */
static void head(orc_conv *c, int slot, int advance, int useX,
                 const float *const *in, float *const *out, const int *run, int nrun)
{
    const int B = c->B, P = c->P;
    int cur = (c->curPart - 1 + P) % P;
    if (g_mutation == 3) cur = (c->curPart + 1) % P;
    if (advance) c->curPart = cur;
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < nrun; r++) {
        int ch = run[r];
        float *d = part_ptr(c->data[ch], c, cur);
        memset(d, 0, sizeof(float) * c->partLen);
        memcpy(d, in[ch], sizeof(float) * B);
        part_r2c(c, d, c->work[ch]);
        float *a = useX ? c->outX[ch] : c->outS[ch];
        planar_cmac(d, part_ptr(c->filt[slot][ch], c, 0), a, B + 8, B + 8);
        part_c2r(c, a, c->work[ch]);
        for (int i = 0; i < B; i++) out[ch][i] = a[i] + c->overlap[ch][i];
        if (advance) memcpy(c->overlap[ch], a + B, sizeof(float) * B);
    }
}

/* ovlNUPProcessTail with M = 1 (ConvolutionImpl.cpp:2539-2643): the sum over
 * partitions 1..P-1 of X_{t+1-i} H_i, walking the ring from curPart. */
static void tail(orc_conv *c, int slot, int useX, const int *run, int nrun)
{
    const int B = c->B, P = c->P;
    const int last = (g_mutation == 2) ? P - 1 : P;
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < nrun; r++) {
        int ch = run[r];
        float *a = useX ? c->outX[ch] : c->outS[ch];
        memset(a, 0, sizeof(float) * c->partLen);
        int cp = c->curPart;
        for (int i = 1; i < last; i++) {
            planar_cmac(part_ptr(c->data[ch], c, cp), part_ptr(c->filt[slot][ch], c, i),
                        a, B + 8, B + 8);
            cp = (cp + 1) % P;
        }
    }
}

/* Second half of IFFT(sum_{i=0}^{P-1} X_{t-i} H_i) with the filter in `slot`
 * after the head of block t: the overlap a switch to that filter needs. */
static void full_overlap(orc_conv *c, int slot, const int *run, int nrun)
{
    const int B = c->B, P = c->P;
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < nrun; r++) {
        int ch = run[r];
        float *a = c->acc[ch];
        memset(a, 0, sizeof(float) * c->partLen);
        int cp = c->curPart;
        for (int i = 1; i < P; i++) {
            cp = (cp + 1) % P;
            planar_cmac(part_ptr(c->data[ch], c, cp), part_ptr(c->filt[slot][ch], c, i),
                        a, B + 8, B + 8);
        }
        planar_cmac(part_ptr(c->data[ch], c, c->curPart), part_ptr(c->filt[slot][ch], c, 0),
                    a, B + 8, B + 8);
        part_c2r(c, a, c->work[ch]);
        memcpy(c->ovNew[ch], a + B, sizeof(float) * B);
    }
}

/* Crossfade, CPU branch (ConvolutionImpl.cpp:1395-1406):
 * out = new * (j/F) + old * ((F-j)/F), step = 1/F in float. */
static void crossfade(orc_conv *c, float *const *out, float *const *newer,
                      const int *run, int nrun)
{
    const int F = c->B;
    const float step = (float)(1.0 / (float)F);
    for (int r = 0; r < nrun; r++) {
        int ch = run[r];
        for (int j = 0; j < c->B; j++) {
            float w1 = step * (float)j, w2 = step * (float)(F - j);
            if (g_mutation == 5) { float t = w1; w1 = w2; w2 = t; }
            out[ch][j] = newer[ch][j] * w1 + out[ch][j] * w2;
        }
    }
}

static int running_list(orc_conv *c, int *run)
{
    int n = 0;
    for (int ch = 0; ch < c->C; ch++)
        if (!c->available[ch]) run[n++] = ch;
    return n;
}

/* ProcessFinalize (ConvolutionImpl.cpp:1046-1079) and the update thread
 * hand-off it triggers (ConvolutionImpl.cpp:2004-2160). */
static void finalize_internal(orc_conv *c)
{
    int *run = (int *)malloc(sizeof(int) * c->C);
    int nrun = running_list(c, run);
    if (nrun > 0) {
        if (c->xmode == ORC_XFADE_REFERENCE) {
            if (c->crossFading) tail(c, c->idxPrev, 0, run, nrun);
            tail(c, c->idxFilter, c->crossFading, run, nrun);
        } else {
            /* intended: the tail for the next block, with both filters when
             * the next block is the crossfade block */
            tail(c, c->idxFilter, 0, run, nrun);
            if (c->delayed) tail(c, c->idxUpdate, 1, run, nrun);
        }
    }
    if (!c->crossFading && c->pending) {
        /* the update thread transforms the staged IR into the update slot;
         * channels not updated keep the current IR (ConvolutionImpl.cpp:2149-2158) */
        for (int ch = 0; ch < c->C; ch++) {
            if (c->stagedSet[ch]) transform_ir(c, c->idxUpdate, ch);
            else memcpy(c->filt[c->idxUpdate][ch], c->filt[c->idxFilter][ch],
                        sizeof(float) * c->bufLen);
            c->stagedSet[ch] = 0;
        }
        c->pending = 0;
        c->readyEvent = 1;
    }
    free(run);
}

int orc_conv_finalize(orc_conv *c)
{
    if (!c) return AMF_NOT_INITIALIZED;
    finalize_internal(c);
    return AMF_OK;
}

int orc_conv_update_td(orc_conv *c, const float *const *ir, size_t n,
                       const uint32_t *flags, uint32_t opflags)
{
    (void)opflags;
    if (!c) return AMF_NOT_INITIALIZED;
    if (!ir) return AMF_INVALID_ARG;
    if (n > (size_t)c->L) return AMF_INVALID_ARG;
    for (int ch = 0; ch < c->C; ch++) {
        if (flags && (flags[ch] & FLAG_FLUSH_STREAM)) flush_channel(c, ch);
        if (flags && (flags[ch] & FLAG_STOP_INPUT)) continue;
        memset(c->staged[ch], 0, sizeof(float) * c->len);
        memcpy(c->staged[ch], ir[ch], sizeof(float) * n);
        c->stagedSet[ch] = 1;
    }
    c->pending = 1;
    return AMF_OK;
}

/* channel flag bookkeeping (ConvolutionImpl.cpp:2868-2907).  A channel that
 * returns from the stopped (available) state starts from silence: the
 * reference flushes a draining channel before its last drain block and would
 * otherwise resume with that block's state. */
static void channel_flags(orc_conv *c, const uint32_t *flags, size_t n)
{
    for (int ch = 0; ch < c->C; ch++) {
        const unsigned char was_available = c->available[ch];
        if (!flags) {
            c->available[ch] = 0;
            c->tailLeft[ch] = c->len;
        } else if (flags[ch] & FLAG_STOP_INPUT) {
            if (flags[ch] & FLAG_FLUSH_STREAM) {
                c->available[ch] = 1;
                c->tailLeft[ch] = 0;
            } else if (c->tailLeft[ch] > 0) {
                c->available[ch] = 0;
                c->tailLeft[ch] -= (int)n;
                if (c->tailLeft[ch] <= 0) flush_channel(c, ch);
            } else {
                c->available[ch] = 1;
                c->tailLeft[ch] = 0;
            }
        } else {
            c->available[ch] = 0;
            c->tailLeft[ch] = c->len;
        }
        if (flags && (flags[ch] & FLAG_FLUSH_STREAM)) {
            if (!c->flushed[ch]) flush_channel(c, ch);
            c->flushed[ch] = 1;
        } else {
            c->flushed[ch] = 0;
        }
        if (was_available && !c->available[ch]) flush_channel(c, ch);
    }
}

int orc_conv_process(orc_conv *c, const float *const *in, float *const *out,
                     size_t n, const uint32_t *flags, size_t *processed)
{
    if (!c) return AMF_NOT_INITIALIZED;
    if (!in || !out) return AMF_INVALID_ARG;
    if (processed) *processed = 0;
    c->lastXfade = 0;
    if (n < (size_t)c->B) return AMF_OK; /* a partial block is not processed */

    channel_flags(c, flags, n);
    int *run = (int *)malloc(sizeof(int) * c->C);
    int nrun = running_list(c, run);
    if (nrun == 0) { free(run); return AMF_WRONG_STATE; }

    /* IR switch schedule (ConvolutionImpl.cpp:869-904): a transformed update
     * arms the crossfade; the following block performs it. */
    int doXfade = 0;
    if (!c->pending && c->readyEvent && !c->delayed) {
        c->readyEvent = 0;
        c->delayed = 1;
    } else {
        doXfade = c->delayed;
    }
    c->crossFading = doXfade;
    c->lastXfade = doXfade;

    if (doXfade) {
        c->idxPrev = c->idxFilter;
        c->idxFilter = (c->idxFilter + 1) % 3;
        c->idxUpdate = (c->idxUpdate + 1) % 3;
        if (c->xmode == ORC_XFADE_REFERENCE) {
            if (c->fmode) {
                head(c, c->idxPrev, 1, 1, in, out, run, nrun);
                head(c, c->idxFilter, 0, 0, in, c->xfade, run, nrun);
            } else {
                head(c, c->idxPrev, 1, 0, in, out, run, nrun);
                finalize_internal(c);
                head(c, c->idxFilter, 1, 0, in, c->xfade, run, nrun);
                finalize_internal(c);
            }
        } else {
            /* old filter: previous tail + overlap; new filter: its own tail
             * and overlap, one FDL advance */
            const int B = c->B, P = c->P;
            int cur = (c->curPart - 1 + P) % P;
            if (g_mutation == 3) cur = (c->curPart + 1) % P;
            c->curPart = cur;
            #pragma omp parallel for schedule(static)
            for (int r = 0; r < nrun; r++) {
                int ch = run[r];
                float *d = part_ptr(c->data[ch], c, cur);
                memset(d, 0, sizeof(float) * c->partLen);
                memcpy(d, in[ch], sizeof(float) * B);
                part_r2c(c, d, c->work[ch]);
                float *ao = c->outS[ch], *an = c->outX[ch];
                planar_cmac(d, part_ptr(c->filt[c->idxPrev][ch], c, 0), ao, B + 8, B + 8);
                planar_cmac(d, part_ptr(c->filt[c->idxFilter][ch], c, 0), an, B + 8, B + 8);
                part_c2r(c, ao, c->work[ch]);
                part_c2r(c, an, c->work[ch]);
                for (int i = 0; i < B; i++) {
                    out[ch][i] = ao[i] + c->overlap[ch][i];
                    c->xfade[ch][i] = an[i] + c->ovNew[ch][i];
                }
                memcpy(c->overlap[ch], an + B, sizeof(float) * B);
            }
            c->delayed = 0;
            if (!c->fmode) finalize_internal(c);
        }
        crossfade(c, out, c->xfade, run, nrun);
        c->delayed = 0;
    } else {
        head(c, c->idxFilter, 1, 0, in, out, run, nrun);
        if (c->delayed && c->xmode == ORC_XFADE_INTENDED)
            full_overlap(c, c->idxUpdate, run, nrun);
        if (!c->fmode) finalize_internal(c);
    }
    free(run);
    if (processed) *processed = (size_t)c->B;
    return AMF_OK;
}

int orc_conv_get_next_free_channel(orc_conv *c, uint32_t *idx, const uint32_t *flags)
{
    if (!c) return AMF_NOT_INITIALIZED;
    if (!idx) return AMF_INVALID_ARG;
    *idx = (uint32_t)-1;
    if (!flags) return AMF_NOT_FOUND;
    for (int ch = 0; ch < c->C; ch++) {
        if ((flags[ch] & FLAG_STOP_INPUT) && c->tailLeft[ch] <= 0) {
            *idx = (uint32_t)ch;
            return AMF_OK;
        }
    }
    return AMF_NOT_FOUND;
}

int orc_conv_last_was_crossfade(const orc_conv *c) { return c ? c->lastXfade : 0; }

/* ---- CPU baseline timing ------------------------------------------------ */

static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double uni01(uint64_t *s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

double orc_bench_upols(int C, int L, int B, int nblocks, int warm, uint64_t seed,
                       double *times)
{
    orc_conv *c = orc_conv_create(L, B, C, ORC_XFADE_INTENDED, 0);
    if (!c) return -1.0;
    uint64_t s = seed;
    float **ir = alloc_rows(C, L);
    for (int ch = 0; ch < C; ch++) {
        double e = 0.0;
        for (int k = 0; k < L; k++) {
            double u1 = uni01(&s), u2 = uni01(&s);
            double g = sqrt(-2.0 * log(u1 + 1e-300)) * cos(2.0 * M_PI * u2);
            double v = g * exp(-(double)k / 12000.0);
            ir[ch][k] = (float)v;
            e += v * v;
        }
        float nrm = (float)(1.0 / sqrt(e));
        for (int k = 0; k < L; k++) ir[ch][k] *= nrm;
    }
    float **x = alloc_rows(C, B), **y = alloc_rows(C, B);
    orc_conv_update_td(c, (const float *const *)ir, (size_t)L, NULL, 0);
    double total = 0.0;
    for (int b = 0; b < warm + nblocks; b++) {
        for (int ch = 0; ch < C; ch++)
            for (int i = 0; i < B; i++) x[ch][i] = (float)(2.0 * uni01(&s) - 1.0);
        double t0 = now_s();
        size_t done = 0;
        orc_conv_process(c, (const float *const *)x, y, (size_t)B, NULL, &done);
        double dt = now_s() - t0;
        if (b >= warm) {
            if (times) times[b - warm] = dt;
            total += dt;
        }
    }
    free_rows(ir, C);
    free_rows(x, C);
    free_rows(y, C);
    orc_conv_destroy(c);
    return total;
}

/* ------------------------------------------------------------------------ */
/* 4. I/O chain                                                              */
/* ------------------------------------------------------------------------ */

void orc_mix(const float *const *in, int channels, size_t n, float *out)
{
    for (size_t i = 0; i < n; i++) {
        float acc = 0.0f;
        for (int c = 0; c < channels; c++) acc = (c == 0) ? in[0][i] : acc + in[c][i];
        out[i] = acc;
    }
}

int orc_convert(const void *in, size_t in_step, size_t in_offset, int in_type, void *out, size_t out_step,
                size_t out_offset, int out_type, size_t n, float gain, int *clipped)
{
    if ((in_type != 0 && in_type != 1) || (out_type != 0 && out_type != 1)) return -1;
    int clip = 0;
    for (size_t i = 0; i < n; i++) {
        const size_t si = i * in_step + in_offset, so = i * out_step + out_offset;
        if (in_type == 0 && out_type == 1) {
            const float scale = 32767.0f * gain;
            float f = ((const float *)in)[si] * scale;
            if (f > 32767.0f) { f = 32767.0f; clip = 1; }
            if (f < -32768.0f) { f = -32768.0f; clip = 1; }
            ((int16_t *)out)[so] = (int16_t)f;  /* C conversion truncates toward zero */
        } else if (in_type == 1 && out_type == 0) {
            const float scale = gain / 32767.0f;
            ((float *)out)[so] = (float)((const int16_t *)in)[si] * scale;
        } else if (in_type == 0) {
            ((float *)out)[so] = ((const float *)in)[si];
        } else {
            ((int16_t *)out)[so] = ((const int16_t *)in)[si];
        }
    }
    if (clipped) *clipped = clip;
    return 0;
}
