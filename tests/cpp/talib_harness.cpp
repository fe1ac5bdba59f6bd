// Restatement of the reference's convolution driver
// (tan/tests/src/TALibTestConvolution/TALibTestConvolution.cpp:66-321) against
// this library's C++ interface (include/TrueAudioNext.h), on the driver's own
// input files, with the check the reference leaves out: every output block
// is compared with the CPU oracle (oracle/tan_oracle.h) driven through the
// same calls.  Also exercises TANFFT, TANMixer and TANConverter through the
// C++ interface.
//
// usage: talib_harness <dir with pulsetrainS_0p5s.wav, testresponse.wav, testresponse2.wav> [out.wav]
//
// Like the driver, it writes the convolution output with WriteWaveFileF
// (tanx_wav_write, 16-bit) and checks the file read back.  The context is
// set up as the driver does (TALibTestConvolution.cpp:166-174: TANCreateContext,
// then InitOpenCL with two queues of the caller's), with hipStream_t for
// cl_command_queue; QueryInterface, the OpenCL-named getters and the device
// ProcessDirect overload are checked too.
#include "TrueAudioNext.h"
#include "tan_mi355x.h"
#include "tan_oracle.h"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace amf;

namespace {

// ReadWaveFile (tan/common/wav.cpp:87-238) through the library
// (tanx_wav_read): planar floats, 16-bit samples scaled by 1/32768.
bool read_wav(const std::string &path, int &channels, std::vector<std::vector<float>> &data)
{
    tanx_wav_info info{};
    if (tanx_wav_read_info(path.c_str(), &info) != TANX_OK || info.channels == 0) return false;
    channels = info.channels;
    data.assign(channels, std::vector<float>(info.samples));
    std::vector<float *> rows(channels);
    for (int c = 0; c < channels; c++) rows[c] = data[c].data();
    return tanx_wav_read(path.c_str(), &info, rows.data(), info.samples) == TANX_OK;
}

int fail(const char *what, int rc)
{
    std::printf("FAIL: %s (AMF_RESULT %d)\n", what, rc);
    return 1;
}

}  // namespace

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : "tests/golden";
    int nch = 0, nres_ch = 0, nres2_ch = 0;
    std::vector<std::vector<float>> samples, resp1, resp2;
    if (!read_wav(dir + "/pulsetrainS_0p5s.wav", nch, samples) || !read_wav(dir + "/testresponse.wav", nres_ch, resp1) ||
        !read_wav(dir + "/testresponse2.wav", nres2_ch, resp2) || nch != nres_ch || nch != nres2_ch) {
        std::printf("FAIL: cannot read the input files in %s\n", dir.c_str());
        return 1;
    }
    const int n_samples = 128;  // TALibTestConvolution.cpp:92
    const amf_uint32 NSamples = (amf_uint32)samples[0].size();
    const amf_uint32 NResSamples = (amf_uint32)resp1[0].size();

    // the driver's sequence: a context, then InitOpenCL(general queue,
    // convolution queue) with queues the caller created (they outlive it)
    TANContextPtr ctx;
    AMF_RESULT r = TANCreateContext(TAN_FULL_VERSION, &ctx);
    if (r != AMF_OK) return fail("TANCreateContext", r);
    hipStream_t general_q = nullptr, conv_q = nullptr;
    if (hipSetDevice(0) != hipSuccess || hipStreamCreate(&general_q) != hipSuccess ||
        hipStreamCreate(&conv_q) != hipSuccess)
        return fail("hipStreamCreate", -1);
    if ((r = ctx->InitOpenCL(general_q, conv_q)) != AMF_OK) return fail("InitOpenCL(queues)", r);
    if (ctx->GetOpenCLGeneralQueue() != general_q || ctx->GetOpenCLConvQueue() != conv_q)
        return fail("GetOpenCLGeneralQueue/GetOpenCLConvQueue", -1);
    if (ctx->InitOpenCL(nullptr, nullptr) != AMF_ALREADY_INITIALIZED || ctx->InitHIP(0) != AMF_ALREADY_INITIALIZED)
        return fail("a second Init", -1);
    // a context on the first one's device, from its HIP context
    {
        TANContextPtr ctx2;
        if ((r = TANCreateContext(TAN_FULL_VERSION, &ctx2)) != AMF_OK) return fail("TANCreateContext", r);
        hipCtx_t hc = ctx->GetOpenCLContext();
        if (!hc) return fail("GetOpenCLContext", -1);
        if ((r = ctx2->InitOpenCL(hc)) != AMF_OK || ctx2->GetHIPDevice() != 0) return fail("InitOpenCL(context)", r);
        if (!ctx2->GetOpenCLConvQueue() || ctx2->GetOpenCLConvQueue() == ctx2->GetOpenCLGeneralQueue())
            return fail("the context's own streams", -1);
    }
    TANConvolutionPtr conv;
    if ((r = TANCreateConvolution(ctx, &conv)) != AMF_OK) return fail("TANCreateConvolution", r);
    // QueryInterface: the object under its IID and the base IID, acquired;
    // no other interface; the smart pointer's conversion uses it
    {
        void *q = nullptr;
        if (conv->QueryInterface(TANConvolution::IID(), &q) != AMF_OK || q != conv.GetPtr())
            return fail("QueryInterface(TANConvolution)", -1);
        static_cast<TANConvolution *>(q)->Release();
        if (conv->QueryInterface(TANInterface::IID(), &q) != AMF_OK || q == nullptr)
            return fail("QueryInterface(base)", -1);
        static_cast<TANConvolution *>(q)->Release();
        if (conv->QueryInterface(TANFFT::IID(), &q) != AMF_NO_INTERFACE || q != nullptr)
            return fail("QueryInterface(TANFFT) on a convolution", -1);
        TANInterfacePtr_T<TANInterface> base(conv.GetPtr());
        TANConvolutionPtr back(base);
        TANFFTPtr none(base);
        if (back.GetPtr() != conv.GetPtr() || none.GetPtr() != nullptr) return fail("TANInterfacePtr_T conversion", -1);
    }
    // the driver's GPU-UN method (TALibTestConvolution.cpp:78-85)
    if ((r = conv->InitGpu(TAN_CONVOLUTION_METHOD_FHT_UNIFORM_PARTITIONED, NResSamples, n_samples, nch)) != AMF_OK)
        return fail("InitGpu", r);
    if (conv->InitCpu(TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_UNIFORM, NResSamples, n_samples, nch) != AMF_NOT_SUPPORTED)
        return fail("InitCpu should be unsupported", -1);

    orc_conv *orc = orc_conv_create((int)NResSamples, n_samples, nch, ORC_XFADE_INTENDED, 0);
    if (!orc) return fail("oracle", -1);

    std::vector<float *> pr1(nch), pr2(nch);
    std::vector<std::vector<float>> out(nch, std::vector<float>(NSamples, 0.f)), oref = out;
    for (int c = 0; c < nch; c++) {
        pr1[c] = resp1[c].data();
        pr2[c] = resp2[c].data();
    }
    amf_uint32 flags[] = {0, 0};
    const amf_uint32 blocking = TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY;
    if ((r = conv->UpdateResponseTD(pr1.data(), NResSamples, nullptr, blocking)) != AMF_OK)
        return fail("UpdateResponseTD", r);
    orc_conv_update_td(orc, pr1.data(), NResSamples, nullptr, blocking);

    const int n_blocks = (int)((NSamples + n_samples - 1) / n_samples);
    const int n_IR_updates = 10;
    const int update_interv = n_blocks / n_IR_updates;
    std::vector<float *> in(nch), o(nch), oo(nch);
    int b = 0, xfades = 0, expected_xfades = 0;
    const int full_blocks = (int)(NSamples / n_samples);
    for (amf_uint32 ns = 0; ns < NSamples; ns += n_samples, b++) {
        if (b % update_interv == 0) {
            float **resp = ((b & 1) == 0) ? pr1.data() : pr2.data();
            if ((r = conv->UpdateResponseTD(resp, NResSamples, nullptr, blocking)) != AMF_OK)
                return fail("UpdateResponseTD", r);
            orc_conv_update_td(orc, resp, NResSamples, nullptr, blocking);
            if (b + 2 < full_blocks) expected_xfades++;  // crossfade two blocks later
        }
        const amf_size actual = (NSamples - ns >= (amf_uint32)n_samples) ? n_samples : NSamples - ns;
        for (int c = 0; c < nch; c++) {
            in[c] = samples[c].data() + ns;
            o[c] = out[c].data() + ns;
            oo[c] = oref[c].data() + ns;
        }
        amf_size done = 0, done_o = 0;
        if ((r = conv->Process(in.data(), o.data(), actual, flags, &done)) != AMF_OK) return fail("Process", r);
        orc_conv_process(orc, in.data(), oo.data(), actual, flags, &done_o);
        if (done != done_o) return fail("processed count differs from the oracle", (int)done);
        xfades += orc_conv_last_was_crossfade(orc);
    }
    double maxerr = 0.0, es = 0.0, ee = 0.0;
    for (int c = 0; c < nch; c++)
        for (amf_uint32 i = 0; i < NSamples; i++) {
            const double d = (double)out[c][i] - (double)oref[c][i];
            maxerr = std::fmax(maxerr, std::fabs(d));
            es += (double)oref[c][i] * oref[c][i];
            ee += d * d;
        }
    const double snr = ee > 0 ? 10.0 * std::log10(es / ee) : 999.0;
    std::printf("convolution: %d blocks, %d crossfades, max|gpu-oracle| = %.3e, SNR = %.1f dB\n", b, xfades, maxerr,
                snr);

    // the driver's WriteWaveFileF of the output (TALibTestConvolution.cpp:290),
    // read back: trunc(clamp(y, -1, 1) * 32767) / 32768 per sample
    bool wav_ok = true;
    if (argc > 2) {
        std::vector<const float *> rows(nch);
        for (int c = 0; c < nch; c++) rows[c] = out[c].data();
        if ((r = (AMF_RESULT)tanx_wav_write(argv[2], 48000, (uint16_t)nch, 16, NSamples, rows.data())) != AMF_OK)
            return fail("tanx_wav_write", r);
        int bch = 0;
        std::vector<std::vector<float>> back;
        wav_ok = read_wav(argv[2], bch, back) && bch == nch && back[0].size() == NSamples;
        for (int c = 0; wav_ok && c < nch; c++)
            for (amf_uint32 i = 0; i < NSamples; i++) {
                double v = out[c][i];
                v = v > 1.0 ? 1.0 : v < -1.0 ? -1.0 : v;
                if (back[c][i] != (float)(int16_t)(v * 32767.0) / 32768.0f) wav_ok = false;
            }
        std::printf("wav: output written and read back %s\n", wav_ok ? "exactly" : "WRONG");
    }

    // TANFFT through the C++ interface: R2C planar then C2R planar round trip
    TANFFTPtr fft;
    if ((r = TANCreateFFT(ctx, &fft)) != AMF_OK) return fail("TANCreateFFT", r);
    if ((r = fft->Init()) != AMF_OK) return fail("TANFFT::Init", r);
    const int log2n = 9, n = 1 << log2n;
    std::vector<float> fin(2 * n, 0.f), fspec(2 * n, 0.f), fback(2 * n, 0.f), ospec(2 * n, 0.f);
    for (int i = 0; i < n; i++) fin[i] = samples[0][i + 2000];
    float *pi[] = {fin.data()}, *ps[] = {fspec.data()}, *pb[] = {fback.data()}, *pos[] = {ospec.data()};
    if ((r = fft->Transform(TAN_FFT_R2C_PLANAR_TRANSFORM_DIRECTION_FORWARD, log2n, 1, pi, ps)) != AMF_OK)
        return fail("Transform R2C planar", r);
    orc_fft_transform(ORC_FFT_R2C_PLANAR_FORWARD, log2n, 1, pi, pos);
    if ((r = fft->Transform(TAN_FFT_C2R_PLANAR_TRANSFORM_DIRECTION_BACKWARD, log2n, 1, ps, pb)) != AMF_OK)
        return fail("Transform C2R planar", r);
    double fe = 0.0, re = 0.0;
    for (int k = 0; k <= n / 2; k++) {
        fe = std::fmax(fe, std::fabs((double)fspec[k] - ospec[k]));
        fe = std::fmax(fe, std::fabs((double)fspec[n / 2 + 8 + k] - ospec[n / 2 + 8 + k]));
    }
    for (int i = 0; i < n; i++) re = std::fmax(re, std::fabs((double)fback[i] - fin[i]));
    std::printf("fft: max|gpu-oracle| spectrum = %.3e, round trip = %.3e\n", fe, re);

    // output chain through the C++ interface (the host overloads of
    // TANMixer::Mix and TANConverter::Convert) on the convolution output:
    // mono mix of all channels, then 16-bit PCM; against the oracle's
    // restatement of Mixer.cl / Converter.cl
    TANMixerPtr mixer;
    TANConverterPtr converter;
    if ((r = TANCreateMixer(ctx, &mixer)) != AMF_OK) return fail("TANCreateMixer", r);
    if ((r = TANCreateConverter(ctx, &converter)) != AMF_OK) return fail("TANCreateConverter", r);
    if ((r = mixer->Init(NSamples, nch)) != AMF_OK) return fail("TANMixer::Init", r);
    if ((r = converter->Init()) != AMF_OK) return fail("TANConverter::Init", r);
    std::vector<float *> outs(nch);
    for (int c = 0; c < nch; c++) outs[c] = out[c].data();
    std::vector<float> mono(NSamples), mono_ref(NSamples);
    if ((r = mixer->Mix(outs.data(), mono.data())) != AMF_OK) return fail("TANMixer::Mix", r);
    orc_mix(outs.data(), nch, NSamples, mono_ref.data());
    std::vector<short> pcm(NSamples), pcm_ref(NSamples);
    bool clipped = false;
    int clipped_ref = 0;
    if ((r = converter->Convert(mono.data(), 1, NSamples, pcm.data(), 1, 0.5f, &clipped)) != AMF_OK)
        return fail("TANConverter::Convert", r);
    orc_convert(mono_ref.data(), 1, 0, 0, pcm_ref.data(), 1, 0, 1, NSamples, 0.5f, &clipped_ref);
    const bool chain_ok = mono == mono_ref && pcm == pcm_ref && clipped == (clipped_ref != 0);
    std::printf("output chain: mix %s, convert %s, clipped %d\n", mono == mono_ref ? "bit-exact" : "DIFFERS",
                pcm == pcm_ref ? "bit-exact" : "DIFFERS", (int)clipped);

    // ProcessDirect, device overload: a TIME_DOMAIN object fed the same IRs,
    // inputs and non-zero ranges as a twin on the host overload; the outputs
    // must be the same bits, over two calls (the history carries over)
    bool direct_ok = false;
    {
        const amf_uint32 tdL = 512, tdB = 128, tdC = 2, tdN = 300;
        TANConvolutionPtr th, td;
        if ((r = TANCreateConvolution(ctx, &th)) != AMF_OK || (r = TANCreateConvolution(ctx, &td)) != AMF_OK)
            return fail("TANCreateConvolution (time domain)", r);
        if ((r = th->InitGpu(TAN_CONVOLUTION_METHOD_TIME_DOMAIN, tdL, tdB, tdC)) != AMF_OK ||
            (r = td->InitGpu(TAN_CONVOLUTION_METHOD_TIME_DOMAIN, tdL, tdB, tdC)) != AMF_OK)
            return fail("InitGpu (time domain)", r);
        std::vector<float> hir(tdC * tdL), hin(tdC * 2 * tdN), hout(tdC * 2 * tdN, 0.f), dres(tdC * 2 * tdN, 0.f);
        for (amf_uint32 i = 0; i < tdC * tdL; i++) hir[i] = resp1[i / tdL][i % tdL];
        for (amf_uint32 i = 0; i < tdC * 2 * tdN; i++) hin[i] = samples[i / (2 * tdN)][3000 + i % (2 * tdN)];
        int nz[] = {0, (int)tdL, 17, 400};
        float *dmem = nullptr;
        const size_t nir = tdC * tdL, nio = tdC * 2 * tdN;
        if (hipMalloc(&dmem, (nir + 2 * nio) * sizeof(float)) != hipSuccess) return fail("hipMalloc", -1);
        float *dir = dmem, *din = dmem + nir, *dout = din + nio;
        if (hipMemcpy(dir, hir.data(), nir * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(din, hin.data(), nio * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return fail("hipMemcpy", -1);
        for (amf_uint32 call = 0; call < 2; call++) {
            float *hi[2], *ho[2], *hr[2];
            tan_hip_mem di[2], dout_m[2], dr[2];
            for (amf_uint32 c = 0; c < tdC; c++) {
                hr[c] = hir.data() + c * tdL;
                hi[c] = hin.data() + c * 2 * tdN + call * tdN;
                ho[c] = hout.data() + c * 2 * tdN + call * tdN;
                dr[c] = reinterpret_cast<tan_hip_mem>(dir + c * tdL);
                di[c] = reinterpret_cast<tan_hip_mem>(din + c * 2 * tdN + call * tdN);
                dout_m[c] = reinterpret_cast<tan_hip_mem>(dout + c * 2 * tdN + call * tdN);
            }
            amf_size dh = 0, dd = 0;
            if ((r = th->ProcessDirect(hr, hi, ho, tdN, &dh, nz)) != AMF_OK) return fail("ProcessDirect (host)", r);
            if ((r = td->ProcessDirect(dr, di, dout_m, tdN, &dd, nz)) != AMF_OK)
                return fail("ProcessDirect (device)", r);
            if (dh != tdN || dd != tdN) return fail("ProcessDirect count", (int)dd);
        }
        if (hipStreamSynchronize(conv_q) != hipSuccess ||
            hipMemcpy(dres.data(), dout, nio * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            return fail("hipMemcpy back", -1);
        (void)hipFree(dmem);
        float peak = 0.f;
        for (float v : hout) peak = std::fmax(peak, std::fabs(v));
        direct_ok = dres == hout && peak > 0.f;
        std::printf("process direct: device overload %s the host overload's output\n",
                    direct_ok ? "equals" : "DIFFERS from");
    }

    orc_conv_destroy(orc);
    ctx->Terminate();  // while conv/fft are still alive, as the driver does
    const bool ok = maxerr <= 1e-5 && snr >= 120.0 && xfades == expected_xfades && fe <= 1e-4 && re <= 1e-5 && chain_ok &&
                    wav_ok && direct_ok;
    std::printf("%s\n", ok ? "PASS" : "FAIL");
    return ok ? 0 : 1;
}
