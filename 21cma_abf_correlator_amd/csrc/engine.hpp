// engine.hpp -- the device engine behind the drop-in ABI and the engine entry points.
//
// One Engine lives on one device and owns what every stage needs there: the FFT
// twiddle tables, the baseline tables, the spectrum buffer of one chunk and, for
// the precision forms, the requantization gains. The stages are plain enqueue
// calls on a caller-given stream; nothing here synchronises the host.
//
// Stages (reference kernels.cu:810-845 does the same work as zero + convert +
// cuFFT + corr_kernel; here it is two hand-written kernels):
//   fengine   : validity mask + int8 -> f32 + R2C FFT (length 2*n_freq, Nyquist
//               dropped) + the form's output conversion, in one kernel
//   xengine   : acc[b][c] += sum over the chunk's frames of Y_j conj(Y_k)
//   scale     : acc *= 1.0f/frames (reference kernels.cu:178-185, 890-891)
//   dump      : integrator -> complex64 [baseline][n_freq] visibilities
// and, for the FP16 and int8 forms,
//   calibrate : per-(input, channel) gains from the mean power of a block of frames
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "corr_cuda.h"

namespace corr {

std::string &last_error();
corr_status fail(corr_status st, const std::string &msg);
corr_status failf(corr_status st, const char *fmt, ...);
corr_status hip_check(hipError_t e, const char *what);

/// Spectrum representation between the F- and X-engine.
enum class Form : int {
    kFP32 = 0,  ///< complex64 spectra, FP32 vector X-engine (the drop-in numerics)
    kFP16 = 1,  ///< FP16 complex spectra with per-(input,channel) gains, FP16 MFMA
    kINT8 = 2,  ///< int8 complex spectra clipped to [-7,7] with gains, int8 MFMA
};

/// Largest station count the drop-in ABI accepts on one device (1024 baselines,
/// reference kernels.cu:108-110, 474-483).
constexpr int kMaxBaselinesAbi = 1024;

/// Longest integration of the int8 form: its int32 integrator adds at most
/// 2 * 7 * 7 = 98 per frame (components clamped to +-7).
constexpr uint64_t kInt8MaxFrames = 2147483647ull / 98ull;

/// Largest input count of the FP16/int8 forms (2*n_station <= 96 Gram rows).
constexpr int kMaxStationMfma = 48;

/// FFT sizes this build transforms: n_freq = 2^k with 512 <= n_freq <= 16384;
/// the multiples of 2048 below 16384 with an odd factor, 6144 (3 x 2048),
/// 10240 (5 x 2048), 12288 (3 x 4096) and 14336 (7 x 2048), in one workgroup's
/// LDS; and every multiple of 2048 from 18432 to 1048576, staged through HBM.
bool fft_size_supported(int n_freq);

/// Twiddle table entries of a supported n_freq: exp(-2 pi i m/n_freq) for
/// m < n_freq, then, for n_freq = M * P with M > 1 (P the largest power of two
/// dividing n_freq, at most 16384), exp(-2 pi i m/P) for m < P (the
/// power-of-two sub-transforms).
int fft_twiddle_elems(int n_freq);

struct EngineShape {
    int n_station = 0;
    int n_freq = 0;
    int chunk_frames = 0;   ///< capacity of the spectrum buffer, in frames
    Form form = Form::kFP32;
};

class Engine {
  public:
    /// Create on the current device. Fails with CORR_ERR_NOT_IMPL for an FFT
    /// length this build does not transform.
    static corr_status create(const EngineShape &shape, Engine **out);
    ~Engine();

    const EngineShape &shape() const { return shape_; }
    /// Input bytes per frame: 2*n_freq int8 samples.
    size_t frame_bytes() const { return 2u * (size_t)shape_.n_freq; }
    int n_baseline() const { return n_baseline_; }

    /// F-engine over `frames` frames of every station.
    ///   d_in   : frame f of station s at d_in + s*station_pitch + f*2*n_freq
    ///   d_mask : optional; bit (mask_bit0 + f) of d_mask + s*mask_pitch, LSB
    ///            first within a byte, set = present; NULL = all present
    corr_status fengine(const int8_t *d_in, size_t station_pitch, const uint8_t *d_mask,
                        size_t mask_pitch, int mask_bit0, int frames, hipStream_t s);

    /// acc[b*n_freq + c] (complex64) += sum over the last fengine's frames.
    corr_status xengine_fp32(float2 *acc, int frames, hipStream_t s);

    /// The form's X-engine: integ[b*n_freq + c] += sum over the last fengine's
    /// frames; integrator element float2 (FP32), float4 = (sum re, sum im,
    /// compensation re, compensation im) (FP16), int2 (int8).
    corr_status xengine(void *integ, int frames, hipStream_t s);

    /// F- and X-engine over `frames` frames (any count, in runs of at most
    /// chunk_frames), accumulated into `integ` as xengine does; same input
    /// arguments as fengine. FP16/int8 forms with more than one run: the
    /// F-engine runs on `s`, the X-engine on the engine's own stream, and the
    /// spectra alternate between two buffers, so the X-engine of one run
    /// overlaps the F-engine of the next. Otherwise (the FP32 form, or one run)
    /// both run in turn on `s`. Every launch is ordered after the work already
    /// enqueued on `s`, and `s` waits for the last X-engine before this
    /// returns: work enqueued on `s` afterwards sees the whole integration. The
    /// first overlapped call allocates the second spectrum buffer, the stream
    /// and its events.
    corr_status integrate(const int8_t *d_in, size_t station_pitch, const uint8_t *d_mask,
                          size_t mask_pitch, int mask_bit0, int64_t frames, void *integ, hipStream_t s);

    /// Allocate what integrate needs (done by its first overlapped call
    /// otherwise): for callers that capture integrate into a graph, where
    /// allocation is barred.
    corr_status prepare_integrate() { return shape_.form == Form::kFP32 ? CORR_OK : ensure_pipeline(); }

    /// Bytes of one integrator element of this form.
    size_t integrator_elem_bytes() const;

    /// Visibilities of an integration of `frames` frames (the caller's nominal
    /// count): out[b*n_freq + c] complex64; FP32 form as the reference's row
    /// scale, FP16/int8 forms rescaled by 1/(g_j g_k frames) in FP64.
    corr_status dump(const void *integ, uint64_t frames, float2 *out, hipStream_t s);

    /// FP16/int8 forms: gains from the mean power of `frames` frames (same
    /// arguments as fengine, any number of frames); overwrites the spectra.
    corr_status calibrate(const int8_t *d_in, size_t station_pitch, const uint8_t *d_mask,
                          size_t mask_pitch, int mask_bit0, int frames, hipStream_t s);

    /// FP16/int8 forms: gains [station][n_freq] (positive, finite) from device memory.
    corr_status set_gains(const float *d_gains, hipStream_t s);
    const float *gains() const { return d_gain_; }

    /// FP16/int8 forms: channels per X-engine workgroup (16 or 32) and the
    /// channel tile of the spectrum buffer's order (the workgroup width or
    /// n_freq, see forms::spec_index). Chosen at creation by timing F- and
    /// X-engine with every combination that fits (once per process and
    /// shape); set_xtile overrides it (before the next fengine call).
    int xtile() const { return x_tile_; }
    int spec_tile() const { return spec_tile_; }
    corr_status set_xtile(int channels, int spec_tile);
    /// Time-K of the matrix-core X-engine: frames per staged step, 1, 2 or 4
    /// K extents of the form's MFMA (frames_per_block(): 16 FP32, 32 FP16,
    /// 64 int8). Chosen by the tuner with the tile; set_xstep overrides it.
    int xstep() const { return x_mfma_ ? x_kb_ * frames_per_block() : 0; }
    corr_status set_xstep(int frames_per_step);
    int frames_per_block() const;

    /// acc *= 1.0f/(float)frames over `n` complex values.
    static corr_status scale(float2 *acc, size_t n, uint64_t frames, hipStream_t s);

    /// Spectrum buffer: FP32 [station][chunk_frames][n_freq] complex64; FP16 /
    /// int8 [station][n_freq/spec_tile][chunk_frames][spec_tile] of the packed
    /// pair (forms::spec_index).
    void *spectra() const { return d_spec_; }
    size_t spectra_station_stride() const { return (size_t)shape_.chunk_frames * shape_.n_freq; }
    size_t spectra_elem_bytes() const;

    /// Device-side launch statistics for the last X-engine plan (for reports).
    int xengine_waves() const { return x_waves_; }

  private:
    Engine() = default;
    corr_status init();

    corr_status fengine_to(void *spec, const int8_t *d_in, size_t station_pitch, const uint8_t *d_mask,
                           size_t mask_pitch, int mask_bit0, int frames, hipStream_t s);
    /// The transform of `frames` frames (validity already expanded, or NULL)
    /// into `out` in representation `form` (0 FP32, 1 FP16, 2 int8; tile of the
    /// spectrum order for 1 and 2): one kernel, or for n_freq above 16384 the
    /// staged kernels in batches of stage_frames_.
    corr_status transform(int form, const int8_t *d_in, size_t station_pitch, const int32_t *valid,
                          size_t valid_stride, int frames, void *out, size_t out_station_stride, int tile,
                          hipStream_t s);
    corr_status xengine_from(const void *spec, void *integ, int frames, hipStream_t s);
    corr_status ensure_pipeline();
    bool overlaps(int64_t frames) const { return shape_.form != Form::kFP32 && frames > shape_.chunk_frames; }

    // integrate(): second spectrum buffer, X-engine stream, per-buffer events.
    void *d_spec_b_ = nullptr;
    hipStream_t xs_ = nullptr;
    hipEvent_t ev_f_[2] = {nullptr, nullptr};  ///< F-engine into buffer i done
    hipEvent_t ev_x_[2] = {nullptr, nullptr};  ///< X-engine out of buffer i done

    EngineShape shape_{};
    int n_baseline_ = 0;

    int cal_frames_ = 1;        ///< frames per calibration batch (FP32 spectra in d_spec_)
    bool x_mfma_ = true;        ///< X-engine on the matrix cores (up to 48 inputs)
    int x_tile_ = 16;           ///< channels per MFMA X-engine workgroup
    int x_kb_ = 1;              ///< MFMA K extents per staged step (time-K)
    int spec_tile_ = 16;        ///< channel tile of the spectrum order (x_tile_ or n_freq)
    corr_status tune_xtile();

    float2 *d_tw_ = nullptr;    ///< fft_twiddle_elems(n_freq) entries
    float2 *d_wl_ = nullptr;    ///< exp(-2 pi i k / (2 n_freq)), k < n_freq
    void *d_spec_ = nullptr;
    int32_t *d_valid_ = nullptr;  ///< expanded validity, [station][chunk_frames]
    /// n_freq above 16384: the staged transform's buffer, [station][frame][n_freq]
    /// complex64 for stage_frames_ frames
    float2 *d_stage_ = nullptr;
    int stage_frames_ = 0;
    int2 *d_pairs_ = nullptr;     ///< (j, k) of every baseline
    float *d_gain_ = nullptr;     ///< FP16/int8 forms: [station][n_freq]
    float *d_power_ = nullptr;    ///< calibration: [station][n_freq] sum of |Y|^2
    int32_t *d_count_ = nullptr;  ///< calibration: frames seen per station

    // FP32 X-engine plan: 4x4 station blocks, lower-triangle block pairs.
    int2 *d_bpairs_ = nullptr;
    int n_bpairs_ = 0;
    int x_npair_ = 1;   ///< block pairs per thread
    int x_waves_ = 1;   ///< waves per workgroup
    int x_groups_ = 1;  ///< workgroups along the pair axis
    int x_frames_lds_ = 4;
};

// Kernel launchers (engine_kernels.hip).
/// 4-bit packed frames (n_freq bytes each, in_pitch bytes per station) to int8
/// frames (2*n_freq bytes, out_pitch per station), `frames` frames of each station.
/// n_freq above 16384: both stages of the staged transform for `frames` frames
/// starting at frame0 (input, validity and output are indexed by frame0 + f;
/// zs holds [station][frames][n_freq] complex64).
hipError_t launch_fengine_staged(int form, int n_freq, const int8_t *in, size_t station_pitch, const int32_t *valid,
                                 size_t valid_stride, int frame0, int frames, int n_station, const float2 *tw,
                                 const float2 *wl, const float *gain, void *out, size_t out_station_stride, int tile,
                                 float2 *zs, hipStream_t s);
hipError_t launch_unpack4(const int8_t *in, size_t in_pitch, int8_t *out, size_t out_pitch, int frames,
                          int n_freq, int n_station, hipStream_t s);
/// tw: fft_twiddle_elems(n_freq) entries; wl: n_freq.
hipError_t launch_twiddles(float2 *tw, float2 *wl, int n_freq, hipStream_t s);
hipError_t launch_validity(const uint8_t *mask, size_t mask_pitch, int bit0, int frames,
                           int n_station, int32_t *valid, size_t valid_stride, hipStream_t s);
/// tile: channels of the spectrum tile, 16, 32 or n_freq (the plain order).
hipError_t launch_fengine_fp32(int n_freq, const int8_t *in, size_t station_pitch,
                               const int32_t *valid, size_t valid_stride, const float2 *tw,
                               const float2 *wl, float2 *out, size_t out_station_stride, int tile,
                               int frames, int n_station, hipStream_t s);
hipError_t launch_xengine_fp32(const float2 *y, size_t y_station_stride, int n_freq,
                               int frames, int n_station, const int2 *bpairs, int n_bpairs,
                               int npair, int waves, int groups, int frames_lds, float2 *acc,
                               hipStream_t s);
hipError_t launch_scale(float2 *acc, size_t n, float s, hipStream_t st);
hipError_t launch_fengine_form(int form, int n_freq, const int8_t *in, size_t station_pitch,
                               const int32_t *valid, size_t valid_stride, const float2 *tw,
                               const float2 *wl, const float *gain, void *out,
                               size_t out_station_stride, int tile, int frames, int n_station,
                               hipStream_t s);
hipError_t launch_power(int nf, int frames, int n_station, const float2 *spec,
                        size_t spec_station_stride, const int32_t *valid, size_t valid_stride,
                        float *power, int32_t *count, hipStream_t s);
hipError_t launch_gains(int form, int nf, int n_station, const float *power, const int32_t *count,
                        float *gain, hipStream_t s);
hipError_t launch_fill(void *p, size_t n_words, uint32_t v, hipStream_t s);
/// n_words of the autotuner's synthetic int8 input (fft::tuning_noise_word).
hipError_t launch_noise(void *p, size_t n_words, hipStream_t s);
/// The matrix-core X-engine of `form` with ch channels per workgroup and kb
/// MFMA K blocks per staged step (xmfma::tile_fits): prepare raises its LDS
/// limit, launch adds `frames` frames of the spectra into the integrator.
hipError_t xengine_mfma_prepare(int form, int n_station, int ch, int kb);
hipError_t launch_xengine_mfma(int form, int n_station, int ch, int kb, const void *spectra, size_t ys, int nf,
                               int spec_tile, int frames, void *integ, hipStream_t s);
hipError_t launch_dump(int form, const void *integ, int n_baseline, int nf, const int2 *pairs,
                       const float *gain, uint64_t frames, float2 *out, hipStream_t s);

}  // namespace corr
