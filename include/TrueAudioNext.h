// TrueAudioNext.h -- C++ interface of the MI355X TAN convolution library.
//
// Source-compatible replacement for the convolution path of TAN's public
// header (tan/tanlibrary/include/TrueAudioNext.h): the same namespace,
// class names, enums, method names and argument order for TANContext
// (:580-597), TANConvolution (:125-250) and TANFFT (:493-541), and the same
// extern "C" factories (:642-681), returning AMF_RESULT codes with the
// reference's numeric values (amf/public/include/core/Result.h:43-122).
//
// Deliberate differences:
//   * OpenCL is replaced by HIP.  Every cl_mem overload is an overload on
//     tan_hip_mem (an opaque pointer type: cast a device float pointer to
//     it); per-channel device arrays must be equally spaced in one
//     allocation.  TANContext::InitOpenCL becomes InitHIP.
//   * No AMF runtime: objects are reference counted (Acquire/Release) and
//     TANInterfacePtr_T is the smart pointer; property storage is not
//     provided.
//   * The CPU engine is not part of this library: InitCpu returns
//     AMF_NOT_SUPPORTED.  Methods TIME_DOMAIN and the non-uniform methods
//     return AMF_NOT_SUPPORTED from Init.
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
    AMF_TAN_CLIPPING_WAS_REQUIRED = 42,
    AMF_TAN_UNSUPPORTED_VERSION = 43
} AMF_RESULT;

// Opaque device-memory handle standing in for cl_mem: a HIP device pointer
// to float data, cast to this type.
typedef struct _tan_hip_mem *tan_hip_mem;

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

// Reference-counted base (the AMFInterface contract: a new object starts at
// count 1; Release at zero deletes it).
class TANInterface {
public:
    virtual amf_long AMF_STD_CALL Acquire() = 0;
    virtual amf_long AMF_STD_CALL Release() = 0;

protected:
    virtual ~TANInterface() {}
};

template <class T>
class TANInterfacePtr_T {
public:
    TANInterfacePtr_T() : p_(nullptr) {}
    TANInterfacePtr_T(T *p) : p_(p) { if (p_) p_->Acquire(); }
    TANInterfacePtr_T(const TANInterfacePtr_T &o) : p_(o.p_) { if (p_) p_->Acquire(); }
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
    virtual AMF_RESULT AMF_STD_CALL Terminate() = 0;
    // replaces InitOpenCL: bind a HIP device (default: device 0 at first use)
    virtual AMF_RESULT AMF_STD_CALL InitHIP(int device = 0) = 0;
    // replaces InitOpenCL(cl_command_queue convolutionQueue, cl_command_queue
    // generalQueue): run on the caller's hipStream_t (nullptr: a stream of the
    // context's own); they must outlive the context
    virtual AMF_RESULT AMF_STD_CALL InitHIPStreams(void *convStream, void *generalStream, int device = 0) = 0;
    // accepted for source compatibility; the GPU path has no OpenMP threads
    virtual AMF_RESULT AMF_STD_CALL InitOpenMP(int nThreads) = 0;
    virtual int AMF_STD_CALL GetHIPDevice() = 0;
    virtual void *AMF_STD_CALL GetHIPGeneralStream() = 0;  // hipStream_t
    virtual void *AMF_STD_CALL GetHIPConvStream() = 0;     // hipStream_t
};
typedef TANInterfacePtr_T<TANContext> TANContextPtr;

class TANConvolution : public TANInterface {
public:
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
    virtual AMF_RESULT AMF_STD_CALL ProcessFinalize(void) = 0;
    virtual AMF_RESULT AMF_STD_CALL GetNextFreeChannel(amf_uint32 *pChannelIndex, const amf_uint32 flagMasks[]) = 0;
};
typedef TANInterfacePtr_T<TANConvolution> TANConvolutionPtr;

class TANFFT : public TANInterface {
public:
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
