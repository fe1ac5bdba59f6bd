// C++ interface (include/TrueAudioNext.h) over the C ABI (include/tan_mi355x.h).
#include "../../include/TrueAudioNext.h"
#include "../../include/tan_mi355x.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <new>
#include <vector>

namespace {

using namespace amf;

template <class Base>
class RefCounted : public Base {
public:
    amf_long AMF_STD_CALL Acquire() override { return ++refs_; }
    amf_long AMF_STD_CALL Release() override
    {
        const amf_long r = --refs_;
        if (r == 0) delete this;
        return r;
    }
    // the object under its interface's IID or the base interface's
    // (AMFInterface::QueryInterface, Interface.h:70), acquired once more
    AMF_RESULT AMF_STD_CALL QueryInterface(const AMFGuid &id, void **pp) override
    {
        if (!pp) return AMF_INVALID_POINTER;
        *pp = nullptr;
        if (id != Base::IID() && id != TANInterface::IID()) return AMF_NO_INTERFACE;
        Acquire();
        *pp = static_cast<Base *>(this);
        return AMF_OK;
    }

protected:
    std::atomic<amf_long> refs_{0};
};

class ContextImpl : public RefCounted<TANContext> {
public:
    // Objects created from the context hold a reference to it, so the HIP
    // context lives until the last of them is released; Terminate only marks
    // the context (callers such as TALibTestConvolution.cpp:260-268 terminate
    // it while their convolution objects are still alive).
    ~ContextImpl() override
    {
        if (ctx_) tanx_ctx_destroy(ctx_);
    }

    AMF_RESULT AMF_STD_CALL Terminate() override
    {
        std::lock_guard<std::mutex> lk(mtx_);
        terminated_ = true;
        return AMF_OK;
    }
    AMF_RESULT AMF_STD_CALL InitHIP(int device) override
    {
        std::lock_guard<std::mutex> lk(mtx_);
        if (ctx_) return AMF_ALREADY_INITIALIZED;
        return (AMF_RESULT)tanx_ctx_create(TAN_FULL_VERSION, device, &ctx_);
    }
    AMF_RESULT AMF_STD_CALL InitHIPStreams(void *convStream, void *generalStream, int device) override
    {
        std::lock_guard<std::mutex> lk(mtx_);
        if (ctx_) return AMF_ALREADY_INITIALIZED;
        return (AMF_RESULT)tanx_ctx_create_on_streams(TAN_FULL_VERSION, device, convStream, generalStream, &ctx_);
    }
    AMF_RESULT AMF_STD_CALL InitOpenCL(hipCtx_t pContext) override
    {
        int device = 0;
        if (pContext) {
            // the device of a HIP context: made current for the query
            // (HIP's context calls are deprecated, not removed)
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wdeprecated-declarations"
            hipCtx_t prev = nullptr;
            if (hipCtxPushCurrent(pContext) != hipSuccess) return AMF_INVALID_ARG;
            const hipError_t e = hipCtxGetDevice(&device);
            (void)hipCtxPopCurrent(&prev);
#pragma clang diagnostic pop
            if (e != hipSuccess) return AMF_INVALID_ARG;
        }
        return InitHIP(device);
    }
    AMF_RESULT AMF_STD_CALL InitOpenCL(hipStream_t general, hipStream_t conv) override
    {
        // the streams' device (the current one for the context's own streams)
        int device = 0;
        hipStream_t any = conv ? conv : general;
        if (any) {
            if (hipStreamGetDevice(any, &device) != hipSuccess) return AMF_INVALID_ARG;
            int other = device;
            if (conv && general && (hipStreamGetDevice(general, &other) != hipSuccess || other != device))
                return AMF_INVALID_ARG;
        } else if (hipGetDevice(&device) != hipSuccess) {
            return AMF_NO_DEVICE;
        }
        return InitHIPStreams(conv, general, device);
    }
    hipCtx_t AMF_STD_CALL GetOpenCLContext() override
    {
        const int d = GetHIPDevice();
        if (d < 0) return nullptr;
        int cur = 0;
        if (hipGetDevice(&cur) != hipSuccess) return nullptr;
        hipCtx_t c = nullptr;
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wdeprecated-declarations"
        if (hipSetDevice(d) == hipSuccess) (void)hipCtxGetCurrent(&c);
#pragma clang diagnostic pop
        (void)hipSetDevice(cur);
        return c;
    }
    hipStream_t AMF_STD_CALL GetOpenCLGeneralQueue() override
    {
        return static_cast<hipStream_t>(GetHIPGeneralStream());
    }
    hipStream_t AMF_STD_CALL GetOpenCLConvQueue() override { return static_cast<hipStream_t>(GetHIPConvStream()); }
    AMF_RESULT AMF_STD_CALL InitOpenMP(int) override { return AMF_OK; }
    int AMF_STD_CALL GetHIPDevice() override
    {
        int32_t d = -1;
        tanx_ctx *c = handle();
        if (c) tanx_ctx_device(c, &d);
        return d;
    }
    // GetOpenCLGeneralQueue / GetOpenCLConvQueue (TrueAudioNext.h:595-596)
    void *AMF_STD_CALL GetHIPGeneralStream() override
    {
        tanx_ctx *c = handle();
        return c ? tanx_ctx_update_stream(c) : nullptr;
    }
    void *AMF_STD_CALL GetHIPConvStream() override
    {
        tanx_ctx *c = handle();
        return c ? tanx_ctx_process_stream(c) : nullptr;
    }

    // the HIP context, created on device 0 at first use when InitHIP was not called
    tanx_ctx *handle()
    {
        std::lock_guard<std::mutex> lk(mtx_);
        if (!ctx_) {
            if (tanx_ctx_create(TAN_FULL_VERSION, 0, &ctx_) != TANX_OK) ctx_ = nullptr;
        }
        return ctx_;
    }

private:
    std::mutex mtx_;
    tanx_ctx *ctx_ = nullptr;
    bool terminated_ = false;
};

// Per-channel device pointers must be one allocation with a constant stride.
bool uniform_stride(const tan_hip_mem *p, amf_uint32 n, size_t *stride_floats)
{
    if (!p || !p[0]) return false;
    if (n == 1) { *stride_floats = 0; return true; }
    const char *b0 = reinterpret_cast<const char *>(p[0]);
    const ptrdiff_t d = reinterpret_cast<const char *>(p[1]) - b0;
    if (d <= 0 || d % sizeof(float)) return false;
    for (amf_uint32 c = 2; c < n; c++)
        if (reinterpret_cast<const char *>(p[c]) - b0 != d * (ptrdiff_t)c) return false;
    *stride_floats = (size_t)d / sizeof(float);
    return true;
}

class ConvolutionImpl : public RefCounted<TANConvolution> {
public:
    explicit ConvolutionImpl(ContextImpl *ctx) : ctx_(ctx) { ctx_->Acquire(); }
    ~ConvolutionImpl() override
    {
        Terminate();
        if (dstage_) (void)hipFree(dstage_);
        ctx_->Release();
    }

    AMF_RESULT AMF_STD_CALL Init(TAN_CONVOLUTION_METHOD m, amf_uint32 L, amf_uint32 B, amf_uint32 C) override
    {
        return InitGpu(m, L, B, C);
    }
    AMF_RESULT AMF_STD_CALL InitCpu(TAN_CONVOLUTION_METHOD, amf_uint32, amf_uint32, amf_uint32) override
    {
        return AMF_NOT_SUPPORTED;
    }
    AMF_RESULT AMF_STD_CALL InitGpu(TAN_CONVOLUTION_METHOD m, amf_uint32 L, amf_uint32 B, amf_uint32 C) override
    {
        if (conv_) return AMF_ALREADY_INITIALIZED;
        tanx_ctx *c = ctx_->handle();
        if (!c) return AMF_NO_DEVICE;
        const AMF_RESULT r = (AMF_RESULT)tanx_conv_create(c, (uint32_t)m, L, B, C, &conv_);
        if (r == AMF_OK) {
            channels_ = C;
            block_ = B;
        }
        return r;
    }
    AMF_RESULT AMF_STD_CALL Terminate() override
    {
        if (conv_) tanx_conv_destroy(conv_);
        conv_ = nullptr;
        return AMF_OK;
    }
    TANContext *AMF_STD_CALL GetContext() override { return ctx_; }

    AMF_RESULT AMF_STD_CALL UpdateResponseTD(float *ppBuffer[], amf_size n, const amf_uint32 flags[],
                                             const amf_uint32 opflags) override
    {
        if (!conv_) return AMF_NOT_INITIALIZED;
        return (AMF_RESULT)tanx_conv_update_td(conv_, ppBuffer, n, flags, opflags);
    }
    AMF_RESULT AMF_STD_CALL UpdateResponseTD(tan_hip_mem ppBuffer[], amf_size n, const amf_uint32 flags[],
                                             const amf_uint32 opflags) override
    {
        if (!conv_) return AMF_NOT_INITIALIZED;
        size_t stride = 0;
        if (!uniform_stride(ppBuffer, channels_, &stride)) return AMF_INVALID_ARG;
        if (channels_ == 1) stride = n;
        return (AMF_RESULT)tanx_conv_update_td_device(conv_, reinterpret_cast<const float *>(ppBuffer[0]), stride,
                                                      n, flags, opflags);
    }
    AMF_RESULT AMF_STD_CALL UpdateResponseFD(float *[], amf_size, const amf_uint32[], const amf_uint32) override
    {
        return conv_ ? AMF_NOT_SUPPORTED : AMF_NOT_INITIALIZED;
    }
    AMF_RESULT AMF_STD_CALL UpdateResponseFD(tan_hip_mem[], amf_size, const amf_uint32[], const amf_uint32) override
    {
        return conv_ ? AMF_NOT_SUPPORTED : AMF_NOT_INITIALIZED;
    }

    AMF_RESULT AMF_STD_CALL Process(float *in[], float *out[], amf_size n, const amf_uint32 flags[],
                                    amf_size *processed) override
    {
        if (!conv_) return AMF_NOT_INITIALIZED;
        size_t done = 0;
        const AMF_RESULT r = (AMF_RESULT)tanx_conv_process(conv_, in, out, n, flags, &done);
        if (processed) *processed = done;
        return r;
    }
    // host input, device output (the reference's Process(float*[], cl_mem[]))
    AMF_RESULT AMF_STD_CALL Process(float *in[], tan_hip_mem out[], amf_size n, const amf_uint32 flags[],
                                    amf_size *processed) override
    {
        if (!conv_) return AMF_NOT_INITIALIZED;
        if (!in) return AMF_INVALID_ARG;
        size_t ostride = 0;
        if (!uniform_stride(out, channels_, &ostride)) return AMF_INVALID_ARG;
        const size_t nb = n / block_, len = nb * block_;
        if (channels_ == 1) ostride = len;
        if (nb == 0) {
            if (processed) *processed = 0;
            return AMF_OK;
        }
        if (stage_floats_ < (size_t)channels_ * len) {
            if (dstage_) (void)hipFree(dstage_);
            dstage_ = nullptr;
            stage_floats_ = 0;
            if (hipMalloc(&dstage_, sizeof(float) * channels_ * len) != hipSuccess) return AMF_OUT_OF_MEMORY;
            stage_floats_ = (size_t)channels_ * len;
        }
        hipStream_t s = static_cast<hipStream_t>(tanx_conv_process_stream(conv_));
        for (amf_uint32 c = 0; c < channels_; c++) {
            if (!in[c]) return AMF_INVALID_ARG;
            if (hipMemcpyAsync(dstage_ + (size_t)c * len, in[c], sizeof(float) * len, hipMemcpyHostToDevice, s) !=
                hipSuccess)
                return AMF_UNEXPECTED;
        }
        size_t done = 0;
        const AMF_RESULT r = (AMF_RESULT)tanx_conv_process_device(
            conv_, dstage_, len, reinterpret_cast<float *>(out[0]), ostride, n, flags, &done);
        if (r == AMF_OK && tanx_conv_synchronize(conv_) != TANX_OK) return AMF_UNEXPECTED;
        if (processed) *processed = done;
        return r;
    }
    AMF_RESULT AMF_STD_CALL Process(tan_hip_mem in[], tan_hip_mem out[], amf_size n, const amf_uint32 flags[],
                                    amf_size *processed) override
    {
        if (!conv_) return AMF_NOT_INITIALIZED;
        size_t istride = 0, ostride = 0;
        if (!uniform_stride(in, channels_, &istride) || !uniform_stride(out, channels_, &ostride))
            return AMF_INVALID_ARG;
        if (channels_ == 1) istride = ostride = (n / block_) * block_;
        size_t done = 0;
        const AMF_RESULT r = (AMF_RESULT)tanx_conv_process_device(
            conv_, reinterpret_cast<const float *>(in[0]), istride, reinterpret_cast<float *>(out[0]), ostride, n,
            flags, &done);
        if (r == AMF_OK && tanx_conv_synchronize(conv_) != TANX_OK) return AMF_UNEXPECTED;
        if (processed) *processed = done;
        return r;
    }
    AMF_RESULT AMF_STD_CALL ProcessDirect(float *ir[], float *in[], float *out[], amf_size n, amf_size *processed,
                                          int *nzFirstLast) override
    {
        // TIME_DOMAIN objects only, as in the reference (ConvolutionImpl.cpp:719-762)
        if (processed) *processed = 0;
        if (!conv_) return AMF_NOT_INITIALIZED;
        size_t done = 0;
        const AMF_RESULT r = (AMF_RESULT)tanx_conv_process_direct(conv_, ir, in, out, n, &done, nzFirstLast);
        if (processed) *processed = done;
        return r;
    }
    AMF_RESULT AMF_STD_CALL ProcessDirect(tan_hip_mem ir[], tan_hip_mem in[], tan_hip_mem out[], amf_size n,
                                          amf_size *processed, int *nzFirstLast) override
    {
        if (processed) *processed = 0;
        if (!conv_) return AMF_NOT_INITIALIZED;
        size_t sh = 0, si = 0, so = 0;
        if (!uniform_stride(ir, channels_, &sh) || !uniform_stride(in, channels_, &si) ||
            !uniform_stride(out, channels_, &so))
            return AMF_INVALID_ARG;
        size_t done = 0;
        const AMF_RESULT r = (AMF_RESULT)tanx_conv_process_direct_device(
            conv_, reinterpret_cast<const float *>(ir[0]), sh, reinterpret_cast<const float *>(in[0]), si,
            reinterpret_cast<float *>(out[0]), so, n, &done, nzFirstLast);
        if (processed) *processed = done;
        return r;
    }
    AMF_RESULT AMF_STD_CALL ProcessFinalize(void) override
    {
        if (!conv_) return AMF_NOT_INITIALIZED;
        return (AMF_RESULT)tanx_conv_finalize(conv_);
    }
    AMF_RESULT AMF_STD_CALL GetNextFreeChannel(amf_uint32 *idx, const amf_uint32 flags[]) override
    {
        if (!conv_) return AMF_NOT_INITIALIZED;
        return (AMF_RESULT)tanx_conv_get_next_free_channel(conv_, idx, flags);
    }

private:
    ContextImpl *ctx_;
    tanx_conv *conv_ = nullptr;
    amf_uint32 channels_ = 0, block_ = 0;
    float *dstage_ = nullptr;
    size_t stage_floats_ = 0;
};

class FFTImpl : public RefCounted<TANFFT> {
public:
    explicit FFTImpl(ContextImpl *ctx) : ctx_(ctx) { ctx_->Acquire(); }
    ~FFTImpl() override { ctx_->Release(); }
    AMF_RESULT AMF_STD_CALL Init() override
    {
        init_ = ctx_->handle() != nullptr;
        return init_ ? AMF_OK : AMF_NO_DEVICE;
    }
    AMF_RESULT AMF_STD_CALL Terminate() override
    {
        init_ = false;
        return AMF_OK;
    }
    TANContext *AMF_STD_CALL GetContext() override { return ctx_; }
    AMF_RESULT AMF_STD_CALL Transform(TAN_FFT_TRANSFORM_DIRECTION dir, amf_uint32 log2len, amf_uint32 channels,
                                      float *in[], float *out[]) override
    {
        if (!init_) return AMF_NOT_INITIALIZED;
        return (AMF_RESULT)tanx_fft_transform(ctx_->handle(), (uint32_t)dir, log2len, channels, in, out);
    }
    AMF_RESULT AMF_STD_CALL Transform(TAN_FFT_TRANSFORM_DIRECTION dir, amf_uint32 log2len, amf_uint32 channels,
                                      tan_hip_mem in[], tan_hip_mem out[]) override
    {
        if (!init_) return AMF_NOT_INITIALIZED;
        size_t si = 0, so = 0;
        if (!uniform_stride(in, channels, &si) || !uniform_stride(out, channels, &so) || si != so)
            return AMF_INVALID_ARG;
        if (channels == 1) si = size_t(2) << log2len;
        return (AMF_RESULT)tanx_fft_transform_device(ctx_->handle(), (uint32_t)dir, log2len, channels,
                                                     reinterpret_cast<const float *>(in[0]),
                                                     reinterpret_cast<float *>(out[0]), si);
    }
    // contiguous batch: channel c at c * dataSpacing floats (0 = 2^(log2len+1))
    AMF_RESULT AMF_STD_CALL TransformBatchGPU(TAN_FFT_TRANSFORM_DIRECTION dir, amf_uint32 log2len,
                                              amf_uint32 channels, tan_hip_mem in, tan_hip_mem out,
                                              int dataSpacing) override
    {
        if (!init_) return AMF_NOT_INITIALIZED;
        const size_t stride = dataSpacing > 0 ? (size_t)dataSpacing : (size_t(2) << log2len);
        return (AMF_RESULT)tanx_fft_transform_device(ctx_->handle(), (uint32_t)dir, log2len, channels,
                                                     reinterpret_cast<const float *>(in),
                                                     reinterpret_cast<float *>(out), stride);
    }

private:
    ContextImpl *ctx_;
    bool init_ = false;
};

class ConverterImpl : public RefCounted<TANConverter> {
public:
    explicit ConverterImpl(ContextImpl *ctx) : ctx_(ctx) { ctx_->Acquire(); }
    ~ConverterImpl() override { ctx_->Release(); }
    AMF_RESULT AMF_STD_CALL Init() override
    {
        init_ = ctx_->handle() != nullptr;
        return init_ ? AMF_OK : AMF_NO_DEVICE;
    }
    AMF_RESULT AMF_STD_CALL Terminate() override
    {
        init_ = false;
        return AMF_OK;
    }
    TANContext *AMF_STD_CALL GetContext() override { return ctx_; }
    AMF_RESULT AMF_STD_CALL Convert(short *in, amf_size inStep, amf_size n, float *out, amf_size outStep,
                                    float gain) override
    {
        if (!init_) return AMF_NOT_INITIALIZED;
        return (AMF_RESULT)tanx_convert(ctx_->handle(), in, inStep, TANX_SAMPLE_SHORT, out, outStep,
                                        TANX_SAMPLE_FLOAT, n, gain, nullptr);
    }
    AMF_RESULT AMF_STD_CALL Convert(float *in, amf_size inStep, amf_size n, short *out, amf_size outStep, float gain,
                                    bool *clipped) override
    {
        if (!init_) return AMF_NOT_INITIALIZED;
        int32_t c = 0;
        const AMF_RESULT r = (AMF_RESULT)tanx_convert(ctx_->handle(), in, inStep, TANX_SAMPLE_FLOAT, out, outStep,
                                                      TANX_SAMPLE_SHORT, n, gain, clipped ? &c : nullptr);
        if (clipped) *clipped = c != 0;
        return r;
    }
    // batch overloads: one conversion per buffer, clipping OR-ed
    // (ConverterImpl.cpp:255-285, 420-456)
    AMF_RESULT AMF_STD_CALL Convert(short **in, amf_size inStep, amf_size n, float **out, amf_size outStep,
                                    float gain, int channels) override
    {
        if (!in || !out || channels <= 0) return AMF_INVALID_ARG;
        for (int i = 0; i < channels; i++) {
            const AMF_RESULT r = Convert(in[i], inStep, n, out[i], outStep, gain);
            if (r != AMF_OK) return r;
        }
        return AMF_OK;
    }
    AMF_RESULT AMF_STD_CALL Convert(float **in, amf_size inStep, amf_size n, short **out, amf_size outStep,
                                    float gain, int channels, bool *clipped) override
    {
        if (!in || !out || channels <= 0) return AMF_INVALID_ARG;
        bool any = false;
        for (int i = 0; i < channels; i++) {
            bool c = false;
            const AMF_RESULT r = Convert(in[i], inStep, n, out[i], outStep, gain, clipped ? &c : nullptr);
            if (r != AMF_OK) return r;
            any = any || c;
        }
        if (clipped) *clipped = any;
        return AMF_OK;
    }
    AMF_RESULT AMF_STD_CALL Convert(tan_hip_mem in, amf_size inStep, amf_size inOff, TAN_SAMPLE_TYPE inType,
                                    tan_hip_mem out, amf_size outStep, amf_size outOff, TAN_SAMPLE_TYPE outType,
                                    amf_size n, float gain, bool *clipped) override
    {
        if (!init_) return AMF_NOT_INITIALIZED;
        int32_t c = 0;
        const AMF_RESULT r = (AMF_RESULT)tanx_convert_device(ctx_->handle(), in, inStep, inOff, (int32_t)inType, out,
                                                             outStep, outOff, (int32_t)outType, n, gain,
                                                             clipped ? &c : nullptr);
        if (clipped) *clipped = c != 0;
        return r;
    }
    AMF_RESULT AMF_STD_CALL Convert(tan_hip_mem *in, amf_size inStep, amf_size *inOff, TAN_SAMPLE_TYPE inType,
                                    tan_hip_mem *out, amf_size outStep, amf_size *outOff, TAN_SAMPLE_TYPE outType,
                                    amf_size n, float gain, int count, bool *clipped) override
    {
        if (!in || !out || !inOff || !outOff || count <= 0) return AMF_INVALID_ARG;
        bool any = false;
        for (int i = 0; i < count; i++) {
            bool c = false;
            const AMF_RESULT r = Convert(in[i], inStep, inOff[i], inType, out[i], outStep, outOff[i], outType, n,
                                         gain, clipped ? &c : nullptr);
            if (r != AMF_OK) return r;
            any = any || c;
        }
        if (clipped) *clipped = any;
        return AMF_OK;
    }

private:
    ContextImpl *ctx_;
    bool init_ = false;
};

class MixerImpl : public RefCounted<TANMixer> {
public:
    explicit MixerImpl(ContextImpl *ctx) : ctx_(ctx) { ctx_->Acquire(); }
    ~MixerImpl() override
    {
        if (m_) tanx_mixer_destroy(m_);
        ctx_->Release();
    }
    AMF_RESULT AMF_STD_CALL Init(amf_size bufferSize, int channels) override
    {
        if (m_) return AMF_ALREADY_INITIALIZED;
        if (!ctx_->handle()) return AMF_NO_DEVICE;
        return (AMF_RESULT)tanx_mixer_create(ctx_->handle(), bufferSize, channels, &m_);
    }
    AMF_RESULT AMF_STD_CALL Terminate() override
    {
        if (m_) tanx_mixer_destroy(m_);
        m_ = nullptr;
        return AMF_OK;
    }
    TANContext *AMF_STD_CALL GetContext() override { return ctx_; }
    AMF_RESULT AMF_STD_CALL Mix(float *in[], float *out) override
    {
        if (!m_) return AMF_NOT_INITIALIZED;
        return (AMF_RESULT)tanx_mix(m_, in, out);
    }
    AMF_RESULT AMF_STD_CALL Mix(tan_hip_mem in[], tan_hip_mem out) override
    {
        if (!m_) return AMF_NOT_INITIALIZED;
        return (AMF_RESULT)tanx_mix_device(m_, reinterpret_cast<const float *const *>(in),
                                           reinterpret_cast<float *>(out));
    }
    AMF_RESULT AMF_STD_CALL Mix(tan_hip_mem in, tan_hip_mem out, amf_size stride) override
    {
        if (!m_) return AMF_NOT_INITIALIZED;
        return (AMF_RESULT)tanx_mix_device_strided(m_, reinterpret_cast<const float *>(in), stride,
                                                   reinterpret_cast<float *>(out));
    }

private:
    ContextImpl *ctx_;
    tanx_mixer *m_ = nullptr;
};

}  // namespace

extern "C" TAN_SDK_LINK AMF_RESULT AMF_CDECL_CALL TANCreateConverter(amf::TANContext *pContext,
                                                                     amf::TANConverter **ppConverter)
{
    if (!pContext || !ppConverter) return AMF_INVALID_POINTER;
    ConverterImpl *c = new (std::nothrow) ConverterImpl(static_cast<ContextImpl *>(pContext));
    if (!c) return AMF_OUT_OF_MEMORY;
    c->Acquire();
    *ppConverter = c;
    return AMF_OK;
}

extern "C" TAN_SDK_LINK AMF_RESULT AMF_CDECL_CALL TANCreateMixer(amf::TANContext *pContext, amf::TANMixer **ppMixer)
{
    if (!pContext || !ppMixer) return AMF_INVALID_POINTER;
    MixerImpl *m = new (std::nothrow) MixerImpl(static_cast<ContextImpl *>(pContext));
    if (!m) return AMF_OUT_OF_MEMORY;
    m->Acquire();
    *ppMixer = m;
    return AMF_OK;
}

extern "C" TAN_SDK_LINK AMF_RESULT AMF_CDECL_CALL TANCreateContext(amf_uint64 version, amf::TANContext **ppContext)
{
    if (!ppContext) return AMF_INVALID_POINTER;
    *ppContext = nullptr;
    if ((version >> 48) != 1) return AMF_TAN_UNSUPPORTED_VERSION;
    ContextImpl *c = new (std::nothrow) ContextImpl();
    if (!c) return AMF_OUT_OF_MEMORY;
    c->Acquire();
    *ppContext = c;
    return AMF_OK;
}

extern "C" TAN_SDK_LINK AMF_RESULT AMF_CDECL_CALL TANCreateConvolution(amf::TANContext *pContext,
                                                                       amf::TANConvolution **ppConvolution)
{
    if (!pContext || !ppConvolution) return AMF_INVALID_POINTER;
    ConvolutionImpl *c = new (std::nothrow) ConvolutionImpl(static_cast<ContextImpl *>(pContext));
    if (!c) return AMF_OUT_OF_MEMORY;
    c->Acquire();
    *ppConvolution = c;
    return AMF_OK;
}

extern "C" TAN_SDK_LINK AMF_RESULT AMF_CDECL_CALL TANCreateFFT(amf::TANContext *pContext, amf::TANFFT **ppFFT)
{
    if (!pContext || !ppFFT) return AMF_INVALID_POINTER;
    FFTImpl *f = new (std::nothrow) FFTImpl(static_cast<ContextImpl *>(pContext));
    if (!f) return AMF_OUT_OF_MEMORY;
    f->Acquire();
    *ppFFT = f;
    return AMF_OK;
}
