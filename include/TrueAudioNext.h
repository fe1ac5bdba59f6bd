// TrueAudioNext.h -- C++ interface of the MI355X TAN convolution library.
//
// Source-compatible replacement for the convolution path of TAN's public
// header (tan/tanlibrary/include/TrueAudioNext.h): the same namespace,
// class names, interface IDs, enums, method names and argument order for
// TANContext (:580-597), TANConvolution (:125-250) and TANFFT (:493-541),
// and the same extern "C" factories (:642-681), returning AMF_RESULT codes
// with the reference's numeric values (amf/public/include/core/Result.h:43-122).
//
// Deliberate differences:
//   * OpenCL is replaced by HIP, under the reference's method names.
//     cl_mem becomes tan_hip_mem (an opaque pointer type: cast a device
//     float pointer to it; per-channel device arrays must be equally spaced
//     in one allocation), cl_command_queue becomes hipStream_t and
//     cl_context becomes hipCtx_t (the context of the device to run on).
//     InitOpenCL / GetOpenCLContext / GetOpenCLGeneralQueue /
//     GetOpenCLConvQueue keep their names and argument order; InitHIP,
//     InitHIPStreams and the GetHIP* getters are the same calls by device
//     ordinal.
//   * No AMF runtime: objects are reference counted (Acquire/Release),
//     QueryInterface answers the reference's interface IDs, and
//     TANInterfacePtr_T is the smart pointer; property storage is not
//     provided.
//   * The CPU engine is not part of this library: InitCpu returns
//     AMF_NOT_SUPPORTED.  Every method, TIME_DOMAIN and the non-uniform
//     ones included, runs on the GPU.
//   * ProcessDirect's device overload takes one device buffer per channel
//     (the reference declares cl_mem* arrays and returns AMF_NOT_SUPPORTED,
//     ConvolutionImpl.cpp:766-778) and runs TIME_DOMAIN objects.
#pragma once

#include <cstddef>
#include <cstdint>

#define TAN_VERSION_MAJOR 1
#define TAN_VERSION_MINOR 3
#define TAN_VERSION_RELEASE 1
#define TAN_VERSION_BUILD 1
#define TAN_FULL_VERSION ((uint64_t(TAN_VERSION_MAJOR) << 48ull) | (uint64_t(TAN_VERSION_MINOR) << 32ull) | \
                          (uint64_t(TAN_VERSION_RELEASE) << 16ull) | uint64_t(TAN_VERSION_BUILD))

#define AMF_STD_CALL
#define AMF_CDECL_CALL
#define TAN_SDK_LINK __attribute__((visibility("default")))

typedef uint64_t amf_uint64;
typedef uint32_t amf_uint32;
typedef uint16_t amf_uint16;
typedef uint8_t amf_uint8;
typedef int64_t amf_int64;
typedef int32_t amf_int32;
typedef size_t amf_size;
typedef long amf_long;

// AMF_RESULT: values as in the reference's Result.h
typedef enum AMF_RESULT {
    AMF_OK = 0,
    AMF_FAIL = 1,
    AMF_UNEXPECTED = 2,
    AMF_ACCESS_DENIED = 3,
    AMF_INVALID_ARG = 4,
    AMF_OUT_OF_RANGE = 5,
    AMF_OUT_OF_MEMORY = 6,
    AMF_INVALID_POINTER = 7,
    AMF_NO_INTERFACE = 8,
    AMF_NOT_IMPLEMENTED = 9,
    AMF_NOT_SUPPORTED = 10,
    AMF_NOT_FOUND = 11,
    AMF_ALREADY_INITIALIZED = 12,
    AMF_NOT_INITIALIZED = 13,
    AMF_INVALID_FORMAT = 14,
    AMF_WRONG_STATE = 15,
    AMF_FILE_NOT_OPEN = 16,
    AMF_NO_DEVICE = 17,
    AMF_INPUT_FULL = 25,
    AMF_TAN_CLIPPING_WAS_REQUIRED = 42,
    AMF_TAN_UNSUPPORTED_VERSION = 43
} AMF_RESULT;

// Opaque device-memory handle standing in for cl_mem: a HIP device pointer
// to float data, cast to this type.
typedef struct _tan_hip_mem *tan_hip_mem;

// HIP's handle types (the same typedefs as hip_runtime_api.h, so this header
// needs no HIP include): hipStream_t stands in for cl_command_queue,
// hipCtx_t for cl_context.
typedef struct ihipStream_t *hipStream_t;
typedef struct ihipCtx_t *hipCtx_t;

namespace amf {

enum TAN_SAMPLE_TYPE { TAN_SAMPLE_TYPE_FLOAT = 0, TAN_SAMPLE_TYPE_SHORT = 1 };

enum TAN_CONVOLUTION_METHOD {
    TAN_CONVOLUTION_METHOD_FFT_OVERLAP_ADD,
    TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_UNIFORM,
    TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM,
    TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED,
    TAN_CONVOLUTION_METHOD_FHT_UNIFORM_PARTITIONED,
    TAN_CONVOLUTION_METHOD_FFT_UNIFORM_HEAD_TAIL,
    TAN_CONVOLUTION_METHOD_FHT_UNIFORM_HEAD_TAIL,
    TAN_CONVOLUTION_METHOD_TIME_DOMAIN,
    TAN_CONVOLUTION_METHOD_FHT_NONUNIFORM_PARTITIONED,
    TAN_CONVOLUTION_METHOD_FFT_NONUNIFORM_PARTITIONED,
    TAN_CONVOLUTION_METHOD_USE_PROCESS_FINALIZE = 0x8000,
    TAN_CONVOLUTION_METHOD_USE_PROCESS_TAILTHREAD = 0xC000,
};

enum TAN_CONVOLUTION_CHANNEL_FLAG {
    TAN_CONVOLUTION_CHANNEL_FLAG_PROCESS = 0,
    TAN_CONVOLUTION_CHANNEL_FLAG_STOP_INPUT = 0x01,
    TAN_CONVOLUTION_CHANNEL_FLAG_FLUSH_STREAM = 0x02,
};

enum TAN_CONVOLUTION_OPERATION_FLAG {
    TAN_CONVOLUTION_OPERATION_FLAG_NONE = 0x00,
    TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY = 0x01,
};

enum TAN_FFT_TRANSFORM_DIRECTION {
    TAN_FFT_TRANSFORM_DIRECTION_FORWARD = 0,
    TAN_FFT_TRANSFORM_DIRECTION_BACKWARD = 1,
    TAN_FFT_R2C_TRANSFORM_DIRECTION_FORWARD = 2,
    TAN_FFT_C2R_TRANSFORM_DIRECTION_BACKWARD = 3,
    TAN_FFT_R2C_PLANAR_TRANSFORM_DIRECTION_FORWARD = 4,
    TAN_FFT_C2R_PLANAR_TRANSFORM_DIRECTION_BACKWARD = 5
};

// Interface ID (the layout of AMF's AMFGuid, Platform.h:387-446).
struct AMFGuid {
    amf_uint32 data1;
    amf_uint16 data2, data3;
    amf_uint8 data41, data42, data43, data44, data45, data46, data47, data48;
    AMFGuid(amf_uint32 d1, amf_uint16 d2, amf_uint16 d3, amf_uint8 a, amf_uint8 b, amf_uint8 c, amf_uint8 d,
            amf_uint8 e, amf_uint8 f, amf_uint8 g, amf_uint8 h)
        : data1(d1), data2(d2), data3(d3), data41(a), data42(b), data43(c), data44(d), data45(e), data46(f),
          data47(g), data48(h)
    {
    }
    bool operator==(const AMFGuid &o) const
    {
        return data1 == o.data1 && data2 == o.data2 && data3 == o.data3 && data41 == o.data41 &&
               data42 == o.data42 && data43 == o.data43 && data44 == o.data44 && data45 == o.data45 &&
               data46 == o.data46 && data47 == o.data47 && data48 == o.data48;
    }
    bool operator!=(const AMFGuid &o) const { return !(*this == o); }
};

// static IID() of an interface, with the reference's value
#define TAN_INTERFACE_ID(d1, d2, d3, a, b, c, d, e, f, g, h)          \
    static const AMFGuid &IID()                                        \
    {                                                                  \
        static const AMFGuid id(d1, d2, d3, a, b, c, d, e, f, g, h);   \
        return id;                                                     \
    }

// Reference-counted base (the AMFInterface contract, Interface.h:63-75: a
// new object starts at count 1; Release at zero deletes it; QueryInterface
// hands out the object, acquired, under its own IID or this one's).
class TANInterface {
public:
    TAN_INTERFACE_ID(0x9d872f34, 0x90dc, 0x4b93, 0xb6, 0xb2, 0x6c, 0xa3, 0x7c, 0x85, 0x25, 0xdb)
    virtual amf_long AMF_STD_CALL Acquire() = 0;
    virtual amf_long AMF_STD_CALL Release() = 0;
    virtual AMF_RESULT AMF_STD_CALL QueryInterface(const AMFGuid &interfaceID, void **ppInterface) = 0;

protected:
    virtual ~TANInterface() {}
};

template <class T>
class TANInterfacePtr_T {
public:
    TANInterfacePtr_T() : p_(nullptr) {}
    TANInterfacePtr_T(T *p) : p_(p) { if (p_) p_->Acquire(); }
    TANInterfacePtr_T(const TANInterfacePtr_T &o) : p_(o.p_) { if (p_) p_->Acquire(); }
    // from a pointer to another interface: QueryInterface for T (null when
    // the object does not implement it)
    template <class U>
    explicit TANInterfacePtr_T(const TANInterfacePtr_T<U> &o) : p_(nullptr)
    {
        void *q = nullptr;
        if (o.GetPtr() && o.GetPtr()->QueryInterface(T::IID(), &q) == AMF_OK) p_ = static_cast<T *>(q);
    }
    ~TANInterfacePtr_T() { if (p_) p_->Release(); }
    TANInterfacePtr_T &operator=(const TANInterfacePtr_T &o)
    {
        if (o.p_) o.p_->Acquire();
        if (p_) p_->Release();
        p_ = o.p_;
        return *this;
    }
    // receives a pointer produced by a factory (already acquired once)
    T **operator&()
    {
        if (p_) { p_->Release(); p_ = nullptr; }
        return &p_;
    }
    T *operator->() const { return p_; }
    operator T *() const { return p_; }
    explicit operator bool() const { return p_ != nullptr; }
    T *GetPtr() const { return p_; }

private:
    T *p_;
};

class TANContext : public TANInterface {
public:
    TAN_INTERFACE_ID(0x601a0b1e, 0x7659, 0x41f2, 0xb2, 0x6d, 0x6a, 0x16, 0x5f, 0x37, 0xa4, 0xa2)
    virtual AMF_RESULT AMF_STD_CALL Terminate() = 0;
    // InitOpenCL(cl_context) (TrueAudioNext.h:587-588): run on the device of
    // a HIP context (the primary context of a device, hipCtxGetCurrent after
    // hipSetDevice; nullptr: device 0), on streams of the context's own
    virtual AMF_RESULT AMF_STD_CALL InitOpenCL(hipCtx_t pContext) = 0;
    // InitOpenCL(general queue, convolution queue) (:589-591): run on the
    // caller's streams (nullptr: a stream of the context's own), on the
    // current device; they must outlive the context
    virtual AMF_RESULT AMF_STD_CALL InitOpenCL(hipStream_t pGeneralQueue = nullptr,
                                               hipStream_t pConvolutionQueue = nullptr) = 0;
    // the same by device ordinal: InitHIP(device) as InitOpenCL(that
    // device's context); InitHIPStreams(conv, general, device) as
    // InitOpenCL(general, conv) on that device
    virtual AMF_RESULT AMF_STD_CALL InitHIP(int device = 0) = 0;
    virtual AMF_RESULT AMF_STD_CALL InitHIPStreams(void *convStream, void *generalStream, int device = 0) = 0;
    // accepted for source compatibility; the GPU path has no OpenMP threads
    virtual AMF_RESULT AMF_STD_CALL InitOpenMP(int nThreads) = 0;
    // GetOpenCLContext / GetOpenCLGeneralQueue / GetOpenCLConvQueue
    // (:593-596): the device's primary context and the context's streams
    virtual hipCtx_t AMF_STD_CALL GetOpenCLContext() = 0;
    virtual hipStream_t AMF_STD_CALL GetOpenCLGeneralQueue() = 0;
    virtual hipStream_t AMF_STD_CALL GetOpenCLConvQueue() = 0;
    virtual int AMF_STD_CALL GetHIPDevice() = 0;
    virtual void *AMF_STD_CALL GetHIPGeneralStream() = 0;  // hipStream_t
    virtual void *AMF_STD_CALL GetHIPConvStream() = 0;     // hipStream_t
};
typedef TANInterfacePtr_T<TANContext> TANContextPtr;

class TANConvolution : public TANInterface {
public:
    TAN_INTERFACE_ID(0x99d46c18, 0x0c92, 0x40bd, 0x8a, 0xc7, 0x1c, 0x30, 0xf7, 0xe4, 0xeb, 0xcb)
    virtual AMF_RESULT AMF_STD_CALL Init(TAN_CONVOLUTION_METHOD convolutionMethod, amf_uint32 responseLengthInSamples,
                                         amf_uint32 bufferSizeInSamples, amf_uint32 channels) = 0;
    virtual AMF_RESULT AMF_STD_CALL InitCpu(TAN_CONVOLUTION_METHOD convolutionMethod,
                                            amf_uint32 responseLengthInSamples, amf_uint32 bufferSizeInSamples,
                                            amf_uint32 channels) = 0;
    virtual AMF_RESULT AMF_STD_CALL InitGpu(TAN_CONVOLUTION_METHOD convolutionMethod,
                                            amf_uint32 responseLengthInSamples, amf_uint32 bufferSizeInSamples,
                                            amf_uint32 channels) = 0;
    virtual AMF_RESULT AMF_STD_CALL Terminate() = 0;
    virtual TANContext *AMF_STD_CALL GetContext() = 0;

    virtual AMF_RESULT AMF_STD_CALL UpdateResponseTD(float *ppBuffer[], amf_size numOfSamplesToProcess,
                                                     const amf_uint32 flagMasks[], const amf_uint32 operationFlags) = 0;
    virtual AMF_RESULT AMF_STD_CALL UpdateResponseTD(tan_hip_mem ppBuffer[], amf_size numOfSamplesToProcess,
                                                     const amf_uint32 flagMasks[], const amf_uint32 operationFlags) = 0;
    virtual AMF_RESULT AMF_STD_CALL UpdateResponseFD(float *ppBuffer[], amf_size numOfSamplesToProcess,
                                                     const amf_uint32 flagMasks[], const amf_uint32 operationFlags) = 0;
    virtual AMF_RESULT AMF_STD_CALL UpdateResponseFD(tan_hip_mem ppBuffer[], amf_size numOfSamplesToProcess,
                                                     const amf_uint32 flagMasks[], const amf_uint32 operationFlags) = 0;

    virtual AMF_RESULT AMF_STD_CALL Process(float *ppBufferInput[], float *ppBufferOutput[],
                                            amf_size numOfSamplesToProcess, const amf_uint32 flagMasks[],
                                            amf_size *pNumOfSamplesProcessed) = 0;
    virtual AMF_RESULT AMF_STD_CALL Process(float *pBufferInput[], tan_hip_mem pBufferOutput[],
                                            amf_size numOfSamplesToProcess, const amf_uint32 flagMasks[],
                                            amf_size *pNumOfSamplesProcessed) = 0;
    virtual AMF_RESULT AMF_STD_CALL Process(tan_hip_mem pBufferInput[], tan_hip_mem pBufferOutput[],
                                            amf_size numOfSamplesToProcess, const amf_uint32 flagMasks[],
                                            amf_size *pNumOfSamplesProcessed) = 0;
    virtual AMF_RESULT AMF_STD_CALL ProcessDirect(float *ppImpulseResponse[], float *ppBufferInput[],
                                                  float *ppBufferOutput[], amf_size numOfSamplesToProcess,
                                                  amf_size *pNumOfSamplesProcessed, int *nzFirstLast = nullptr) = 0;
    // device buffers, one per channel (equally spaced in one allocation
    // each); enqueued on the convolution stream, outputs written when it
    // reaches them
    virtual AMF_RESULT AMF_STD_CALL ProcessDirect(tan_hip_mem ppImpulseResponse[], tan_hip_mem ppBufferInput[],
                                                  tan_hip_mem ppBufferOutput[], amf_size numOfSamplesToProcess,
                                                  amf_size *pNumOfSamplesProcessed, int *nzFirstLast = nullptr) = 0;
    virtual AMF_RESULT AMF_STD_CALL ProcessFinalize(void) = 0;
    virtual AMF_RESULT AMF_STD_CALL GetNextFreeChannel(amf_uint32 *pChannelIndex, const amf_uint32 flagMasks[]) = 0;
};
typedef TANInterfacePtr_T<TANConvolution> TANConvolutionPtr;

class TANFFT : public TANInterface {
public:
    TAN_INTERFACE_ID(0x5f3965a1, 0x46c2, 0x4987, 0xb7, 0x1f, 0x3f, 0xf8, 0x1a, 0xe7, 0x5b, 0x6f)
    virtual AMF_RESULT AMF_STD_CALL Init() = 0;
    virtual AMF_RESULT AMF_STD_CALL Terminate() = 0;
    virtual TANContext *AMF_STD_CALL GetContext() = 0;
    virtual AMF_RESULT AMF_STD_CALL Transform(TAN_FFT_TRANSFORM_DIRECTION direction, amf_uint32 log2len,
                                              amf_uint32 channels, float *pBufferInput[], float *pBufferOutput[]) = 0;
    virtual AMF_RESULT AMF_STD_CALL Transform(TAN_FFT_TRANSFORM_DIRECTION direction, amf_uint32 log2len,
                                              amf_uint32 channels, tan_hip_mem pBufferInput[],
                                              tan_hip_mem pBufferOutput[]) = 0;
    virtual AMF_RESULT AMF_STD_CALL TransformBatchGPU(TAN_FFT_TRANSFORM_DIRECTION direction, amf_uint32 log2len,
                                                      amf_uint32 channels, tan_hip_mem pBufferInput,
                                                      tan_hip_mem pBufferOutput, int dataSpacing) = 0;
};
typedef TANInterfacePtr_T<TANFFT> TANFFTPtr;

// TANConverter (TrueAudioNext.h:344-398 of the reference): conversions run
// on the GPU; host overloads stage through device memory.
class TANConverter : public TANInterface {
public:
    TAN_INTERFACE_ID(0x2c3e6c65, 0x1788, 0x4bbf, 0xb0, 0x70, 0xa1, 0x35, 0x8c, 0xbe, 0xe8, 0x9e)
    virtual AMF_RESULT AMF_STD_CALL Init() = 0;
    virtual AMF_RESULT AMF_STD_CALL Terminate() = 0;
    virtual TANContext *AMF_STD_CALL GetContext() = 0;
    virtual AMF_RESULT AMF_STD_CALL Convert(short *inputBuffer, amf_size inputStep, amf_size numOfSamplesToProcess,
                                            float *outputBuffer, amf_size outputStep, float conversionGain) = 0;
    virtual AMF_RESULT AMF_STD_CALL Convert(float *inputBuffer, amf_size inputStep, amf_size numOfSamplesToProcess,
                                            short *outputBuffer, amf_size outputStep, float conversionGain,
                                            bool *outputClipped = nullptr) = 0;
    virtual AMF_RESULT AMF_STD_CALL Convert(short **inputBuffers, amf_size inputStep, amf_size numOfSamplesToProcess,
                                            float **outputBuffers, amf_size outputStep, float conversionGain,
                                            int channels) = 0;
    virtual AMF_RESULT AMF_STD_CALL Convert(float **inputBuffers, amf_size inputStep, amf_size numOfSamplesToProcess,
                                            short **outputBuffers, amf_size outputStep, float conversionGain,
                                            int channels, bool *outputClipped = nullptr) = 0;
    virtual AMF_RESULT AMF_STD_CALL Convert(tan_hip_mem inputBuffer, amf_size inputStep, amf_size inputOffset,
                                            TAN_SAMPLE_TYPE inputType, tan_hip_mem outputBuffer, amf_size outputStep,
                                            amf_size outputOffset, TAN_SAMPLE_TYPE outputType,
                                            amf_size numOfSamplesToProcess, float conversionGain,
                                            bool *outputClipped = nullptr) = 0;
    virtual AMF_RESULT AMF_STD_CALL Convert(tan_hip_mem *inputBuffers, amf_size inputStep, amf_size *inputOffsets,
                                            TAN_SAMPLE_TYPE inputType, tan_hip_mem *outputBuffers,
                                            amf_size outputStep, amf_size *outputOffsets, TAN_SAMPLE_TYPE outputType,
                                            amf_size numOfSamplesToProcess, float conversionGain, int count,
                                            bool *outputClipped = nullptr) = 0;
};
typedef TANInterfacePtr_T<TANConverter> TANConverterPtr;

// TANMixer (TrueAudioNext.h:611-635 of the reference).
class TANMixer : public TANInterface {
public:
    TAN_INTERFACE_ID(0xd2432f7f, 0xc646, 0x4764, 0x9d, 0x61, 0x60, 0xd2, 0x7c, 0x7e, 0x20, 0xc7)
    virtual AMF_RESULT AMF_STD_CALL Init(amf_size buffer_size, int num_channels) = 0;
    virtual AMF_RESULT AMF_STD_CALL Terminate() = 0;
    virtual TANContext *AMF_STD_CALL GetContext() = 0;
    virtual AMF_RESULT AMF_STD_CALL Mix(float *ppBufferInput[], float *ppBufferOutput) = 0;
    virtual AMF_RESULT AMF_STD_CALL Mix(tan_hip_mem pBufferInput[], tan_hip_mem pBufferOutput) = 0;
    virtual AMF_RESULT AMF_STD_CALL Mix(tan_hip_mem pBufferInput, tan_hip_mem pBufferOutput,
                                        amf_size inputStride) = 0;
};
typedef TANInterfacePtr_T<TANMixer> TANMixerPtr;

}  // namespace amf

extern "C" {
TAN_SDK_LINK AMF_RESULT AMF_CDECL_CALL TANCreateConverter(amf::TANContext *pContext, amf::TANConverter **ppConverter);
TAN_SDK_LINK AMF_RESULT AMF_CDECL_CALL TANCreateMixer(amf::TANContext *pContext, amf::TANMixer **ppMixer);
TAN_SDK_LINK AMF_RESULT AMF_CDECL_CALL TANCreateContext(amf_uint64 version, amf::TANContext **ppContext);
TAN_SDK_LINK AMF_RESULT AMF_CDECL_CALL TANCreateConvolution(amf::TANContext *pContext,
                                                            amf::TANConvolution **ppConvolution);
TAN_SDK_LINK AMF_RESULT AMF_CDECL_CALL TANCreateFFT(amf::TANContext *pContext, amf::TANFFT **ppFFT);
}
