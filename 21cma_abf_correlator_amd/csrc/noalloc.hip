// noalloc.hip -- corrx_correlate_noalloc: one FP32-form integration (F-engine,
// matrix-core X-engine, row scale) that allocates nothing.
//
// It is the work of corrx_integrate + corrx_dump on an FP32 engine (the
// reference's per-chunk device work, corr-cuda/src/kernels.cu:810-845, and row
// scale, :878-903), for callers that cannot call hipMalloc: code under HIP
// graph capture, tight serving loops, harnesses that forbid allocation. In
// place of an engine's buffers it uses `scratch`, a file-scope device array of
// kScratchBytes that every GPU the library runs on holds until the process
// ends:
//   [twiddles: fft_twiddle_elems(n_freq) complex64][R2C twiddles: n_freq]
//   [validity: n_station x K int32][spectra: n_station x K x n_freq complex64,
//    channel tiles of 16]
// each region aligned to 256 bytes, with K the frames per run that fit. The
// caller's output doubles as the integrator: zeroed by a fill kernel, then
// F-engine + X-engine per run of K frames, then scaled by 1.0f/dump_frames in
// place. Every launch goes on the caller's stream; there is no memset node and
// no host synchronisation, so the call can be captured into a graph.
//
// The array's name is a host-side stand-in, not a device address: on its first
// call on a device the entry point looks the address up with
// hipGetSymbolAddress (with that device current) and keeps it per device; the
// kernels only ever receive that address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "corrx.h"
#include "engine.hpp"

#ifndef CORR_HD
#define CORR_HD __host__ __device__
#endif
#include "xmfma_body.hpp"

namespace corr {
namespace {

/// Size of the static scratch array on every device (256 MiB).
constexpr size_t kScratchBytes = size_t(256) << 20;
constexpr size_t kAlign = 256;
/// Channels per X-engine workgroup and spectrum tile of this entry point.
constexpr int kTile = 16;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;

/// Per device: the array's address, looked up once with that device current,
/// and the X-engine kernels' LDS limit raised once.
struct DeviceScratch {
    std::once_flag once;
    void *ptr = nullptr;
    hipError_t err = hipSuccess;
};
DeviceScratch g_scratch[kMaxDevices];

constexpr size_t align_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

hipError_t scratch_of_current_device(void **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    DeviceScratch &d = g_scratch[dev];
    std::call_once(d.once, [&d] {
        d.err = hipGetSymbolAddress(&d.ptr, HIP_SYMBOL(scratch));
        // the FP32 X-engine at 16 channels, every row-tile count (inputs 8, 16,
        // ..., 48): raises the dynamic LDS limit where a count needs it
        for (int n = 8; n <= kMaxStationMfma && d.err == hipSuccess; n += 8)
            d.err = xengine_mfma_prepare(0, n, kTile, 1);
    });
    *out = d.ptr;
    return d.err;
}

}  // namespace
}  // namespace corr

using corr::fail;
using corr::failf;
using corr::hip_check;

extern "C" corr_status corrx_correlate_noalloc(int32_t n_station, int32_t n_freq, const int8_t *d_in,
                                               size_t station_pitch, const uint8_t *d_mask, size_t mask_pitch,
                                               int32_t mask_bit0, int64_t frames, uint64_t dump_frames,
                                               void *d_out, void *stream) {
    const int N = n_station, nf = n_freq;
    if (N < 1 || N > corr::kMaxStationMfma)
        return failf(CORR_ERR_ARG, "correlate_noalloc: n_station=%d outside [1, %d]", N, corr::kMaxStationMfma);
    if (nf > 16384 || !corr::fft_size_supported(nf))
        return failf(CORR_ERR_NOT_IMPL,
                     "correlate_noalloc: n_freq=%d; this entry point transforms n_freq = 2^k in [512, 16384], "
                     "6144, 10240, 12288 and 14336",
                     nf);
    if (d_in == nullptr || d_out == nullptr) return fail(CORR_ERR_ARG, "correlate_noalloc: NULL buffer");
    if (frames < 0) return fail(CORR_ERR_ARG, "correlate_noalloc: negative frame count");
    if (dump_frames == 0) return fail(CORR_ERR_ARG, "correlate_noalloc: dump_frames must be > 0");
    if (station_pitch / (2u * (size_t)nf) < (uint64_t)frames)
        return fail(CORR_ERR_ARG, "correlate_noalloc: station pitch shorter than the frames it holds");
    if (mask_bit0 < 0) return fail(CORR_ERR_ARG, "correlate_noalloc: negative mask bit offset");
    if (d_mask != nullptr && ((uint64_t)mask_bit0 + (uint64_t)frames > (uint64_t)INT32_MAX ||
                              mask_pitch * 8u < (uint64_t)mask_bit0 + (uint64_t)frames))
        return fail(CORR_ERR_ARG, "correlate_noalloc: mask pitch shorter than the bits it must hold");

    // Scratch geometry: tables first, then K frames of validity and spectra.
    const size_t tw_bytes = corr::align_up(sizeof(float2) * (size_t)corr::fft_twiddle_elems(nf));
    const size_t wl_bytes = corr::align_up(sizeof(float2) * (size_t)nf);
    const size_t per_frame = (size_t)N * (sizeof(float2) * (size_t)nf + sizeof(int32_t));
    const size_t room = corr::kScratchBytes - tw_bytes - wl_bytes - 2 * corr::kAlign;
    if (room / per_frame < 1)
        return failf(CORR_ERR_ARG,
                     "correlate_noalloc: one frame of %d inputs at n_freq %d needs %zu bytes of scratch; the "
                     "static scratch holds %zu",
                     N, nf, per_frame + tw_bytes + wl_bytes, corr::kScratchBytes);
    const int64_t fit = (int64_t)(room / per_frame);
    const int K = (int)(frames < fit ? (frames > 0 ? frames : 1) : fit);

    void *base = nullptr;
    corr_status st = hip_check(corr::scratch_of_current_device(&base), "hipGetSymbolAddress(scratch)");
    if (st != CORR_OK) return st;
    unsigned char *p = static_cast<unsigned char *>(base);
    float2 *tw = reinterpret_cast<float2 *>(p);
    float2 *wl = reinterpret_cast<float2 *>(p + tw_bytes);
    int32_t *valid = reinterpret_cast<int32_t *>(p + tw_bytes + wl_bytes);
    float2 *spec = reinterpret_cast<float2 *>(p + tw_bytes + wl_bytes +
                                              corr::align_up(sizeof(int32_t) * (size_t)N * (size_t)K));

    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t B = (size_t)N * (size_t)(N + 1) / 2;
    const size_t spec_stride = (size_t)K * (size_t)nf;
    if ((st = hip_check(corr::launch_fill(d_out, B * (size_t)nf * 2u, 0u, s), "fill_kernel(output)")) != CORR_OK ||
        (st = hip_check(corr::launch_twiddles(tw, wl, nf, s), "twiddle_kernel")) != CORR_OK)
        return st;
    const size_t frame_bytes = 2u * (size_t)nf;
    for (int64_t f0 = 0; f0 < frames; f0 += K) {
        const int n = frames - f0 < K ? (int)(frames - f0) : K;
        const int32_t *v = nullptr;
        if (d_mask != nullptr) {
            if ((st = hip_check(corr::launch_validity(d_mask, mask_pitch, mask_bit0 + (int)f0, n, N, valid, (size_t)K, s),
                                "validity_kernel")) != CORR_OK)
                return st;
            v = valid;
        }
        if ((st = hip_check(corr::launch_fengine_fp32(nf, d_in + (size_t)f0 * frame_bytes, station_pitch, v, (size_t)K,
                                                      tw, wl, spec, spec_stride, corr::kTile, n, N, s),
                            "fengine_fp32_kernel")) != CORR_OK ||
            (st = hip_check(corr::launch_xengine_mfma(0, N, corr::kTile, 1, spec, spec_stride, nf, corr::kTile, n, d_out, s),
                            "xengine_mfma_kernel")) != CORR_OK)
            return st;
    }
    return hip_check(corr::launch_scale(static_cast<float2 *>(d_out), B * (size_t)nf, 1.0f / (float)dump_frames, s),
                     "scale_kernel");
}
