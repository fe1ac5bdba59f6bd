/* CPU oracle, part 5: room impulse responses by image sources.
 *
 * TEST INFRASTRUCTURE ONLY (a checker for the library's
 * tanx_room_response*): a scalar restatement of TrueAudioVR's GPU path, the
 * host set-up of TrueAudioVRimpl::generateRoomResponse
 * (samples/src/TrueAudioVR/TrueAudioVR.cpp:659-778, rotMtx :614-651) and
 * the kernels GenerateRoomResponse (Kernels/GenerateRoomResponse.cl:
 * 204-325) and Fill (Kernels/Fill.cl:30-40), plus the head filters of
 * generateSimpleHeadRelatedTransform (TrueAudioVR.cpp:361-415).
 *
 * Built with -ffp-contract=off and without FMA code generation: every
 * operation rounds to single precision in the reference's order, so the
 * fixed-point image sums are reproducible bit for bit.  Differences from
 * the reference, shared with the library and stated there: integer powers
 * by binary powering (native_powr), saturating float-to-integer conversion
 * with NaN -> 0 (convert_int), signed read-back of the sums (Fill.cl reads
 * them unsigned), and the optional direct-path suppression of the CPU path.
 */
#include "tan_oracle.h"

#include <math.h>
#include <string.h>
#include <stdlib.h>

#define ORC_F2I 67108864.0f /* Float2Int, GenerateRoomResponse.cl:31 */

static float powi_f(float b, int e)
{
    float r = 1.0f;
    while (e > 0) {
        if (e & 1) r = r * b;
        b = b * b;
        e >>= 1;
    }
    return r;
}

static int f2i_sat(float v)
{
    if (v != v) return 0;
    if (v >= 2147483647.0f) return 2147483647;
    if (v <= -2147483648.0f) return -2147483647 - 1;
    return (int)v;
}

int orc_room_ear_params(const orc_room *room, const orc_source *source, const orc_listener *ears, int fs, int len,
                        unsigned flags, int max_bounces, orc_room_ear out[2], long long *images)
{
    const float S = 340.0f;
    float damp[6];
    float sx, sy, sz, dxl, dxr, vxl, vzl, vxr, vzr;
    float v[4][3];
    float m[3][3];
    float yaw, pitch, roll, sY, cY, sP, cP, sR, cR;
    const float pi = (float)3.141592653589793;
    int nsw, nsh, nsl, n[3], ch, k;

    if (!(room->width > 0.0f) || !(room->height > 0.0f) || !(room->length > 0.0f) || fs <= 0 || len <= 0)
        return 4;
    damp[0] = room->damp_right;
    damp[1] = room->damp_left;
    damp[2] = room->damp_top;
    damp[3] = room->damp_bottom;
    damp[4] = room->damp_front;
    damp[5] = room->damp_back;
    for (k = 0; k < 6; k++)
        if (!(damp[k] >= 0.0f && damp[k] <= 1.0f)) return 4;
    if (ears->filter_length < 0 || ears->filter_length > ORC_HEAD_FILTER) return 4;

    /* TrueAudioVR.cpp:663-668: the speaker is kept inside the room */
    sx = source->x;
    sy = source->y;
    sz = source->z;
    if (sx > room->width) sx = room->width;
    if (sx < 0) sx = 0;
    if (sy > room->height) sy = room->height;
    if (sy < 0) sy = 0;
    if (sz > room->length) sz = room->length;
    if (sz < 0) sz = 0;

    /* :671-688 ear offsets and directions, rotated */
    dxl = ears->ear_spacing / 2;
    dxr = -ears->ear_spacing / 2;
    vxl = dxl;
    vzl = -dxl;
    vxr = dxr;
    vzr = dxr;
    v[0][0] = dxl; v[0][1] = 0.0f; v[0][2] = 0.0f;
    v[1][0] = dxr; v[1][1] = 0.0f; v[1][2] = 0.0f;
    v[2][0] = vxl; v[2][1] = 0.0f; v[2][2] = vzl;
    v[3][0] = vxr; v[3][1] = 0.0f; v[3][2] = vzr;
    yaw = ears->yaw;
    pitch = ears->pitch;
    roll = ears->roll;
    sY = sinf(yaw * pi / 180);
    cY = cosf(yaw * pi / 180);
    sP = sinf(pitch * pi / 180);
    cP = cosf(pitch * pi / 180);
    sR = sinf(roll * pi / 180);
    cR = cosf(roll * pi / 180);
    m[0][0] = cR * cY - sR * sP * sY;
    m[0][1] = -sR * cP;
    m[0][2] = cR * sY + sR * sP * cY;
    m[1][0] = sR * cY + cR * sP * sY;
    m[1][1] = cR * cP;
    m[1][2] = sR * sY - cR * sP * cY;
    m[2][0] = -cP * sY;
    m[2][1] = sP;
    m[2][2] = cP * cY;
    for (k = 0; k < 4; k++) {
        float x = v[k][0], y = v[k][1], z = v[k][2];
        v[k][0] = x * m[0][0] + y * m[0][1] + z * m[0][2];
        v[k][1] = x * m[1][0] + y * m[1][1] + z * m[1][2];
        v[k][2] = x * m[2][0] + y * m[2][1] + z * m[2][2];
    }

    /* :690-702 images per axis */
    nsw = (int)((room->width / S) * fs);
    nsh = (int)((room->height / S) * fs);
    nsl = (int)((room->length / S) * fs);
    if (nsw < 1 || nsh < 1 || nsl < 1) return 4;
    n[0] = 1 + len / nsw;
    n[1] = 1 + len / nsh;
    n[2] = 1 + len / nsl;
    if (flags & ORC_ROOM_LIMIT_BOUNCES) {
        if (max_bounces < 0) return 4;
        for (k = 0; k < 3; k++)
            if (n[k] > max_bounces) n[k] = max_bounces;
    }
    if ((double)(2 * n[0] + 1) * (double)(2 * n[1] + 1) * (double)(2 * n[2] + 1) > 2147483647.0) return 5;
    *images = (long long)(2 * n[0] + 1) * (2 * n[1] + 1) * (2 * n[2] + 1);

    for (ch = 0; ch < 2; ch++) {
        orc_room_ear *E = &out[ch];
        memset(E, 0, sizeof(*E));
        E->src[0] = sx;
        E->src[1] = sy;
        E->src[2] = sz;
        E->dim[0] = room->width;
        E->dim[1] = room->height;
        E->dim[2] = room->length;
        E->head[0] = ears->head_x + v[ch][0];
        E->head[1] = ears->head_y + v[ch][1];
        E->head[2] = ears->head_z + v[ch][2];
        for (k = 0; k < 3; k++) {
            E->ev[k] = v[2 + ch][k];
            E->n[k] = n[k];
        }
        E->ev_norm = sqrtf(E->ev[0] * E->ev[0] + E->ev[1] * E->ev[1] + E->ev[2] * E->ev[2]);
        for (k = 0; k < 6; k++) E->damp[k] = damp[k];
        E->max_gain = 2.0f;
        E->d_min = 2 * ears->ear_spacing;
        E->fs = fs;
        E->len = len;
        E->hrtf_len = 2 * (nsw + nsh + nsl); /* :752 */
        E->filt_len = ears->filter_length;
        E->suppress_direct = (flags & ORC_ROOM_SUPPRESS_DIRECT) ? 1 : 0;
        for (k = 0; k < ears->filter_length; k++) {
            E->lp[k] = ears->low_pass[k];
            E->hp[k] = ears->high_pass[k];
        }
    }
    return 0;
}

int orc_room_image(const orc_room_ear *E, int ix, int iy, int iz, int *index, int *taps, float *scale, float *hf)
{
    int idx[3] = {ix, iy, iz};
    float pos[3];
    int far_w[3], near_w[3];
    int k, q;
    float amp, dx, dy, dz, d, dr, dp, cosa;

    *index = -1;
    *taps = 0;
    *scale = 0.0f;
    *hf = 0.0f;
    if (E->suppress_direct && ix == 0 && iy == 0 && iz == 0) return 0;
    for (k = 0; k < 3; k++) {
        int i = idx[k];
        /* GenerateRoomResponse.cl:215-234 */
        float a = (i % 2 == 0) ? (float)i : (float)(i + 1);
        float b = (i % 2 == 0) ? 1.0f : -1.0f;
        pos[k] = a * E->dim[k] + b * E->src[k];
        /* :238-244 reflections off the near (left, bottom, back) and far
         * (right, top, front) walls */
        near_w[k] = (i > 0) ? i / 2 : abs(i - 1) / 2;
        far_w[k] = (i > 0) ? (i + 1) / 2 : abs(i) / 2;
    }
    /* :246-251 right, left, top, bottom, front, back */
    amp = powi_f(E->damp[0], far_w[0]) * powi_f(E->damp[1], near_w[0]) * powi_f(E->damp[2], far_w[1]) *
          powi_f(E->damp[3], near_w[1]) * powi_f(E->damp[4], far_w[2]) * powi_f(E->damp[5], near_w[2]);
    dx = pos[0] - E->head[0];
    dy = pos[1] - E->head[1];
    dz = pos[2] - E->head[2];
    d = sqrtf(dx * dx + dy * dy + dz * dz);
    q = f2i_sat((d / 340.0f) * (float)E->fs);
    if (q < 0 || q >= E->len - 1) return 0;
    *index = q + 1;
    dr = (d <= E->d_min) ? E->max_gain : E->max_gain * E->d_min / d;
    k = E->len - *index;
    *taps = E->filt_len > k ? k : E->filt_len;
    dp = E->ev[0] * dx + E->ev[1] * dy + E->ev[2] * dz;
    cosa = dp / (E->ev_norm * d);
    *hf = (cosa + 1.0f) / 2.0f;
    *scale = ORC_F2I * amp * dr;
    return *index < E->hrtf_len ? 2 : 1; /* 2: head-filtered taps, 1: one tap */
}

int orc_room_response(const orc_room *room, const orc_source *source, const orc_listener *listener, int fs,
                      int len, unsigned flags, int max_bounces, float *left, float *right)
{
    orc_room_ear ears[2];
    long long images = 0;
    int rc = orc_room_ear_params(room, source, listener, fs, len, flags, max_bounces, ears, &images);
    unsigned *acc;
    int ch;
    if (rc) return rc;
    acc = (unsigned *)calloc((size_t)len, sizeof(unsigned));
    if (!acc) return 6;
    for (ch = 0; ch < 2; ch++) {
        const orc_room_ear *E = &ears[ch];
        float *out = ch ? right : left;
        int ix, iy, iz, i;
        memset(acc, 0, (size_t)len * sizeof(unsigned));
        for (iz = -E->n[2]; iz <= E->n[2]; iz++)
            for (iy = -E->n[1]; iy <= E->n[1]; iy++)
                for (ix = -E->n[0]; ix <= E->n[0]; ix++) {
                    int index, taps, kind;
                    float s, hf;
                    kind = orc_room_image(E, ix, iy, iz, &index, &taps, &s, &hf);
                    if (kind == 1) {
                        acc[index] += (unsigned)f2i_sat(s);
                    } else if (kind == 2) {
                        for (i = 0; i < taps; i++)
                            acc[index + i] += (unsigned)f2i_sat(s * (E->lp[i] + hf * E->hp[i]));
                    }
                }
        /* Fill.cl:37 (signed) */
        for (i = 0; i < len; i++) out[i] = (float)(int)acc[i] / ORC_F2I;
    }
    free(acc);
    return 0;
}

int orc_room_head_model(float ear_spacing, orc_listener *listener)
{
    /* TrueAudioVR.cpp:361-415: FFT of an impulse (all ones), scaled by
     * fc / (fc + f_j), inverse FFT (1/N), real part */
    enum { N = ORC_HEAD_FILTER };
    double spec[2 * N], imp[2 * N], t[2 * N];
    float fc, in_delta;
    int j;
    if (!(ear_spacing > 0.0f)) return 4;
    fc = (float)(343.0 / (1.10 * ear_spacing));
    in_delta = (float)(1.0 / (float)48000);
    memset(imp, 0, sizeof(imp));
    imp[0] = 1.0;
    orc_dft_f64(N, imp, spec, 0);
    for (j = 0; j < N; j++) {
        float f = (float)((float)(j / 2) / ((float)N * in_delta));
        float d = (float)(fc / (fc + f));
        spec[2 * j] *= d;
        spec[2 * j + 1] *= d;
    }
    orc_dft_f64(N, spec, t, 1);
    listener->filter_length = N;
    for (j = 0; j < N; j++) listener->low_pass[j] = (float)(t[2 * j] / N);
    listener->high_pass[0] = 1.0f;
    for (j = 1; j < N; j++) listener->high_pass[j] = 0.0f;
    for (j = 0; j < N; j++) listener->high_pass[j] -= listener->low_pass[j];
    return 0;
}
