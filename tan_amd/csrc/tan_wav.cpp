// WAV file I/O for impulse responses and test signals: the behaviour of the
// reference's reader and writers (tan/common/wav.cpp) behind the C ABI.
//
// Reading follows ReadWaveFile (wav.cpp:87-238): the RIFF chunk walk to the
// "fmt " and "data" chunks, samples = data bytes * 8 / (bits * channels),
// and the sample scaling 8-bit (u - 127) / 256, 16-bit s / 32768, 32-bit
// the stored IEEE float.  Writing follows SetupWaveHeader + WriteWaveFileF /
// WriteWaveFileS (wav.cpp:29-87, 240-401): a 44-byte canonical header,
// format tag 3 for 32-bit samples and 1 otherwise, and float samples clamped
// to [-1, 1] and converted toward zero after scaling by 127 (8-bit, written
// as signed bytes) or 32767 (16-bit).
//
// Where the reference's walk would loop forever or read past the end of a
// damaged file (a missing "data" chunk, a truncated header), these calls
// return AMF_INVALID_FORMAT instead; samples missing from a truncated data
// chunk read as zero.  Skipped chunks of odd length are followed by their
// RIFF pad byte, which the reference's walk does not skip.
#include "tan_mi355x.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

namespace {

struct FileCloser {
    void operator()(FILE *f) const
    {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

bool read_exact(FILE *f, void *p, size_t n) { return std::fread(p, 1, n, f) == n; }

uint32_t le32(const unsigned char *p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }
uint16_t le16(const unsigned char *p) { return (uint16_t)(p[0] | p[1] << 8); }

// Skips a chunk body of `len` bytes and its pad byte.
bool skip_chunk(FILE *f, uint32_t len) { return std::fseek(f, (long)len + (long)(len & 1), SEEK_CUR) == 0; }

// Header walk of ReadWaveFile (wav.cpp:107-171).  Leaves the file at the
// first sample byte and returns the data chunk's byte length in *data_bytes.
int32_t parse_header(FILE *f, tanx_wav_info *info, uint64_t *data_bytes)
{
    unsigned char riff[12];
    if (!read_exact(f, riff, 8) || std::memcmp(riff, "RIFF", 4) != 0) return TANX_INVALID_FORMAT;
    unsigned char id[4], len[4];
    // chunks up to the "WAVE" form type (wav.cpp:119-131)
    for (;;) {
        if (!read_exact(f, id, 4)) return TANX_INVALID_FORMAT;
        if (std::memcmp(id, "WAVE", 4) == 0) break;
        if (!read_exact(f, len, 4) || !skip_chunk(f, le32(len))) return TANX_INVALID_FORMAT;
    }
    // chunks up to "fmt " (wav.cpp:133-146)
    for (;;) {
        if (!read_exact(f, id, 4)) return TANX_INVALID_FORMAT;
        if (std::memcmp(id, "fmt ", 4) == 0) break;
        if (!read_exact(f, len, 4) || !skip_chunk(f, le32(len))) return TANX_INVALID_FORMAT;
    }
    unsigned char fmt[20];
    if (!read_exact(f, fmt, 20)) return TANX_INVALID_FORMAT;
    const uint32_t fmt_len = le32(fmt);
    if (fmt_len < 16) return TANX_INVALID_FORMAT;
    info->format_tag = le16(fmt + 4);
    info->channels = le16(fmt + 6);
    info->sample_rate = le32(fmt + 8);
    info->bits_per_sample = le16(fmt + 18);
    info->reserved = 0;
    if (fmt_len > 16 && !skip_chunk(f, fmt_len - 16)) return TANX_INVALID_FORMAT;  // same parity as fmt_len
    // chunks up to "data" (wav.cpp:152-158)
    for (;;) {
        if (!read_exact(f, id, 4) || !read_exact(f, len, 4)) return TANX_INVALID_FORMAT;
        if (std::memcmp(id, "data", 4) == 0) break;
        if (!skip_chunk(f, le32(len))) return TANX_INVALID_FORMAT;
    }
    if (info->channels == 0) return TANX_INVALID_FORMAT;
    const int bits = info->bits_per_sample;
    if (bits != 8 && bits != 16 && bits != 32) return TANX_INVALID_FORMAT;
    *data_bytes = le32(len);
    info->samples = (uint32_t)((*data_bytes * 8) / ((uint64_t)bits * info->channels));
    return TANX_OK;
}

int32_t open_and_parse(const char *path, File &f, tanx_wav_info *info, uint64_t *data_bytes)
{
    if (!path || !info) return TANX_INVALID_POINTER;
    f.reset(std::fopen(path, "rb"));
    if (!f) return TANX_FILE_NOT_OPEN;
    return parse_header(f.get(), info, data_bytes);
}

// Interleaved sample bytes of the data chunk, zero past a truncated end.
std::vector<unsigned char> read_samples(FILE *f, const tanx_wav_info &info)
{
    const size_t bytes = (size_t)info.samples * info.channels * (info.bits_per_sample / 8);
    std::vector<unsigned char> raw(bytes, 0);
    if (bytes) (void)std::fread(raw.data(), 1, bytes, f);
    return raw;
}

// SetupWaveHeader (wav.cpp:29-87): RIFF, WAVE, "fmt " (16 bytes), "data".
void put_header(unsigned char *h, uint32_t rate, uint16_t bits, uint16_t channels, uint32_t samples)
{
    auto put32 = [](unsigned char *p, uint32_t v) {
        for (int i = 0; i < 4; i++) p[i] = (unsigned char)(v >> (8 * i));
    };
    auto put16 = [](unsigned char *p, uint16_t v) {
        p[0] = (unsigned char)v;
        p[1] = (unsigned char)(v >> 8);
    };
    const uint32_t data_len = samples * channels * (bits / 8);
    std::memcpy(h, "RIFF", 4);
    put32(h + 4, data_len + 36);  // data + the 36 header bytes after the RIFF length
    std::memcpy(h + 8, "WAVEfmt ", 8);
    put32(h + 16, 16);
    put16(h + 20, bits == 32 ? 3 : 1);
    put16(h + 22, channels);
    put32(h + 24, rate);
    put32(h + 28, (uint32_t)channels * rate * (bits / 8));
    put16(h + 32, (uint16_t)(channels * (bits / 8)));
    put16(h + 34, bits);
    std::memcpy(h + 36, "data", 4);
    put32(h + 40, data_len);
}

}  // namespace

extern "C" {

int32_t tanx_wav_read_info(const char *path, tanx_wav_info *info)
{
    File f;
    uint64_t bytes = 0;
    return open_and_parse(path, f, info, &bytes);
}

int32_t tanx_wav_read(const char *path, tanx_wav_info *info, float *const *out, size_t capacity)
{
    File f;
    uint64_t bytes = 0;
    int32_t rc = open_and_parse(path, f, info, &bytes);
    if (rc != TANX_OK) return rc;
    if (!out) return TANX_INVALID_POINTER;
    for (uint32_t c = 0; c < info->channels; c++)
        if (!out[c]) return TANX_INVALID_POINTER;
    const std::vector<unsigned char> raw = read_samples(f.get(), *info);
    const size_t n = info->samples < capacity ? info->samples : capacity;
    const uint32_t C = info->channels;
    for (uint32_t c = 0; c < C; c++) {
        float *o = out[c];
        switch (info->bits_per_sample) {
        case 8:  // wav.cpp:205-213
            for (size_t i = 0; i < n; i++) o[i] = (float)((int)raw[i * C + c] - 127) / 256.0f;
            break;
        case 16:  // wav.cpp:214-223
            for (size_t i = 0; i < n; i++) {
                const int16_t s = (int16_t)le16(raw.data() + 2 * (i * C + c));
                o[i] = (float)s / 32768.0f;
            }
            break;
        default:  // 32: wav.cpp:224-233
            for (size_t i = 0; i < n; i++) std::memcpy(&o[i], raw.data() + 4 * (i * C + c), 4);
            break;
        }
    }
    return TANX_OK;
}

int32_t tanx_wav_read_raw(const char *path, tanx_wav_info *info, void *data, size_t capacity_bytes)
{
    File f;
    uint64_t bytes = 0;
    int32_t rc = open_and_parse(path, f, info, &bytes);
    if (rc != TANX_OK) return rc;
    if (!data) return TANX_INVALID_POINTER;
    const std::vector<unsigned char> raw = read_samples(f.get(), *info);
    if (raw.size() > capacity_bytes) return TANX_OUT_OF_RANGE;
    std::memcpy(data, raw.data(), raw.size());
    return TANX_OK;
}

int32_t tanx_wav_write(const char *path, uint32_t sample_rate, uint16_t channels, uint16_t bits, uint32_t samples,
                       const float *const *in)
{
    if (!path || !in) return TANX_INVALID_POINTER;
    if (channels == 0 || (bits != 8 && bits != 16 && bits != 32)) return TANX_INVALID_ARG;
    for (uint16_t c = 0; c < channels; c++)
        if (!in[c]) return TANX_INVALID_POINTER;
    const size_t width = bits / 8;
    std::vector<unsigned char> buf(44 + (size_t)samples * channels * width);
    put_header(buf.data(), sample_rate, bits, channels, samples);
    unsigned char *p = buf.data() + 44;
    for (size_t i = 0; i < samples; i++) {
        for (uint16_t c = 0; c < channels; c++, p += width) {
            float v = in[c][i];
            if (bits == 32) {  // wav.cpp:300-308
                std::memcpy(p, &v, 4);
                continue;
            }
            // wav.cpp:268-299: clamp, scale in double, convert toward zero
            if (v > 1.0f) v = 1.0f;
            if (v < -1.0f) v = -1.0f;
            if (bits == 8) {
                p[0] = (unsigned char)(signed char)((double)v * 127.0);
            } else {
                const int16_t s = (int16_t)((double)v * 32767.0);
                p[0] = (unsigned char)(s & 0xff);
                p[1] = (unsigned char)((uint16_t)s >> 8);
            }
        }
    }
    File f(std::fopen(path, "wb"));
    if (!f) return TANX_FILE_NOT_OPEN;
    if (std::fwrite(buf.data(), 1, buf.size(), f.get()) != buf.size()) return TANX_FAIL;
    return TANX_OK;
}

int32_t tanx_wav_write_short(const char *path, uint32_t sample_rate, uint16_t channels, uint16_t bits,
                             uint32_t samples, const int16_t *interleaved)
{
    if (!path || !interleaved) return TANX_INVALID_POINTER;
    if (channels == 0) return TANX_INVALID_ARG;
    if (bits != 16) return TANX_NOT_IMPLEMENTED;  // wav.cpp:343-355
    unsigned char h[44];
    put_header(h, sample_rate, bits, channels, samples);
    File f(std::fopen(path, "wb"));
    if (!f) return TANX_FILE_NOT_OPEN;
    const size_t n = (size_t)samples * channels;
    std::vector<unsigned char> body(2 * n);
    for (size_t i = 0; i < n; i++) {
        body[2 * i] = (unsigned char)(interleaved[i] & 0xff);
        body[2 * i + 1] = (unsigned char)((uint16_t)interleaved[i] >> 8);
    }
    if (std::fwrite(h, 1, 44, f.get()) != 44 || std::fwrite(body.data(), 1, body.size(), f.get()) != body.size())
        return TANX_FAIL;
    return TANX_OK;
}

}  // extern "C"
