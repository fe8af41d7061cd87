// engine_kernels.hip -- gfx950 kernels of the F-engine and the FP32 X-engine.
//
// F-engine (replaces reference kernels.cu:130-139 int8_to_float, :159-174
// zero_invalid_frames and the cuFFT R2C of :578-592/:832-835):
//   One workgroup transforms one frame of one station. The 2*nf real samples are
//   read as nf complex values z[m] = x[2m] + i x[2m+1], transformed by an nf-point
//   complex FFT held in LDS, and unpacked into the nf retained R2C bins
//     X[k] = (Z[k] + conj(Z[nf-k]))/2 + W_L^k (Z[k] - conj(Z[nf-k]))/(2i),
//   W_L = exp(-2 pi i / (2 nf)), k = 0..nf-1 (the Nyquist bin is never formed).
//   The complex FFT is a Stockham autosort FFT (Stockham 1966, in the radix-R
//   GPU formulation of Govindaraju et al., SC'08): radix-16 passes with a last
//   radix-2/4/8 pass, each pass a register-resident radix-R DFT (radix-2 DIF
//   inside the registers) between LDS exchanges. Twiddles come from tables built
//   once per engine in FP64. A frame whose validity bit is clear is never read:
//   the workgroup transforms zeros, which is bit-for-bit what a host-zeroed
//   frame gives.
//
// FP32 X-engine (replaces reference kernels.cu:194-265 corr_kernel):
//   acc[b][c] += sum_t Y_j[t][c] conj(Y_k[t][c]), Re = xr*yr + xi*yi,
//   Im = xi*yr - xr*yi (kernels.cu:243-246), summed in frame order per chunk and
//   added into the row once per chunk. A workgroup owns 64 channels (one per
//   lane); each wave owns 4x4 blocks of station pairs, so a thread keeps 16
//   complex accumulators per block and does 64 FMAs per 8 LDS reads.
#include "engine.hpp"

#ifndef CORR_HD
#define CORR_HD __host__ __device__
#endif
#include "fft_body.hpp"
#include "xengine_body.hpp"
#include "xmfma_body.hpp"

namespace corr {
namespace {

struct DeviceSync {
    __device__ void operator()() const { __syncthreads(); }
};

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef int intx4_t __attribute__((ext_vector_type(4)));

// The wave-wide operations of the MFMA X-engine on gfx950.
struct MfmaOps {
    // D = A B + C, 16x16x32 FP16 -> FP32
    __device__ xmfma::Acc<float> mma(const xmfma::Frag &a, const xmfma::Frag &b,
                                     const xmfma::Acc<float> &c) const {
        const floatx4_t d = __builtin_amdgcn_mfma_f32_16x16x32_f16(
            __builtin_bit_cast(half8_t, a), __builtin_bit_cast(half8_t, b),
            __builtin_bit_cast(floatx4_t, c), 0, 0, 0);
        return __builtin_bit_cast(xmfma::Acc<float>, d);
    }
    // D = A B + C, 16x16x64 int8 -> int32
    __device__ xmfma::Acc<int32_t> mma(const xmfma::Frag &a, const xmfma::Frag &b,
                                       const xmfma::Acc<int32_t> &c) const {
        const intx4_t d = __builtin_amdgcn_mfma_i32_16x16x64_i8(
            __builtin_bit_cast(intx4_t, a), __builtin_bit_cast(intx4_t, b),
            __builtin_bit_cast(intx4_t, c), 0, 0, 0);
        return __builtin_bit_cast(xmfma::Acc<int32_t>, d);
    }
    // D = A B + C as four 16x16x4 FP32 -> FP32 products in turn, the i-th
    // taking word i of each lane's fragments (one float of A and of B per lane)
    __device__ xmfma::Acc<float> mma_f32x4(const xmfma::Frag &a, const xmfma::Frag &b,
                                           const xmfma::Acc<float> &c) const {
        floatx4_t d = __builtin_bit_cast(floatx4_t, c);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            d = __builtin_amdgcn_mfma_f32_16x16x4f32(__builtin_bit_cast(float, a.w[i]),
                                                     __builtin_bit_cast(float, b.w[i]), d, 0, 0, 0);
        return __builtin_bit_cast(xmfma::Acc<float>, d);
    }
    // value of lane l^1
    __device__ float xor1(float x) const { return __shfl_xor(x, 1); }
    __device__ int32_t xor1(int32_t x) const { return __shfl_xor(x, 1); }
};

// ---------------------------------------------------------------------------
// F-engine, complex64 output
// grid (frames, stations), block N/16 threads, N = n_freq = 2^LOG2N
// ---------------------------------------------------------------------------
// At n_freq 16384 the transform exchanges through a half buffer (69.6 KB of
// LDS) so that two 1024-lane workgroups share a CU: 8 waves per SIMD, 64
// registers per lane.
template <int LOG2N>
__global__ void __launch_bounds__((1 << LOG2N) / 16)
__attribute__((amdgpu_waves_per_eu(LOG2N == 14 ? 8 : 1)))
fengine_fp32_kernel(const int8_t *__restrict__ in, size_t station_pitch,
                    const int32_t *__restrict__ valid, size_t valid_stride,
                    const float2 *__restrict__ tw, const float2 *__restrict__ wl,
                    float2 *__restrict__ out, size_t out_station_stride, int tile_shift) {
    __shared__ float2 lds[fft::fengine_lds_elems<LOG2N>()];
    DeviceSync sync;
    fft::fengine_fp32_block<LOG2N>((int)threadIdx.x, (int)blockIdx.x, (int)blockIdx.y, lds, sync, in,
                                   station_pitch, valid, valid_stride, tw, wl, out,
                                   out_station_stride, tile_shift);
}

// F-engine of the FP16 (FORM 1) and int8 (FORM 2) forms: as above, then gain
// and conversion per bin; output [station][frame][N] packed complex.
template <int LOG2N, int FORM>
__global__ void __launch_bounds__((1 << LOG2N) / 16)
__attribute__((amdgpu_waves_per_eu(LOG2N == 14 ? 8 : 1)))
fengine_form_kernel(const int8_t *__restrict__ in, size_t station_pitch,
                    const int32_t *__restrict__ valid, size_t valid_stride,
                    const float2 *__restrict__ tw, const float2 *__restrict__ wl,
                    const float *__restrict__ gain, void *__restrict__ out,
                    size_t out_station_stride, int tile_shift) {
    __shared__ float2 lds[fft::fengine_lds_elems<LOG2N>()];
    DeviceSync sync;
    fft::fengine_form_block<LOG2N, FORM>((int)threadIdx.x, (int)blockIdx.x, (int)blockIdx.y, lds,
                                         sync, in, station_pitch, valid, valid_stride, tw, wl, gain,
                                         out, out_station_stride, tile_shift);
}

// F-engine at n_freq = M * 2^LOG2P (M = 3, 5, 7): the radix-M stage, then M
// power-of-two sub-transforms side by side in LDS (fft_body.hpp); one
// workgroup per (frame, station), N/16 lanes, up to 122 KB of LDS.
template <int M, int LOG2P>
__global__ void __launch_bounds__((M << LOG2P) / 16)
fengine_mixed_fp32_kernel(const int8_t *__restrict__ in, size_t station_pitch,
                          const int32_t *__restrict__ valid, size_t valid_stride,
                          const float2 *__restrict__ tw, const float2 *__restrict__ wl,
                          float2 *__restrict__ out, size_t out_station_stride, int tile) {
    __shared__ float2 lds[fft::MixedSizes<M, LOG2P>::kLds];
    DeviceSync sync;
    fft::fengine_mixed_fp32_block<M, LOG2P>((int)threadIdx.x, (int)blockIdx.x, (int)blockIdx.y, lds, sync, in,
                                            station_pitch, valid, valid_stride, tw, wl, out, out_station_stride,
                                            tile);
}

template <int M, int LOG2P, int FORM>
__global__ void __launch_bounds__((M << LOG2P) / 16)
fengine_mixed_form_kernel(const int8_t *__restrict__ in, size_t station_pitch,
                          const int32_t *__restrict__ valid, size_t valid_stride,
                          const float2 *__restrict__ tw, const float2 *__restrict__ wl,
                          const float *__restrict__ gain, void *__restrict__ out,
                          size_t out_station_stride, int tile) {
    __shared__ float2 lds[fft::MixedSizes<M, LOG2P>::kLds];
    DeviceSync sync;
    fft::fengine_mixed_form_block<M, LOG2P, FORM>((int)threadIdx.x, (int)blockIdx.x, (int)blockIdx.y, lds, sync,
                                                  in, station_pitch, valid, valid_stride, tw, wl, gain, out,
                                                  out_station_stride, tile);
}

// n_freq above 16384, the two-kernel form (fft_body.hpp staged_*). Stage 1: grid (m, frames, stations), P/16 lanes, the P-point
// exchange buffer (139 KB at P = 16384).
template <int LOG2P>
__global__ void __launch_bounds__((1 << LOG2P) / 16)
staged_first_sub_kernel(const int8_t *__restrict__ in, size_t station_pitch, const int32_t *__restrict__ valid,
                        size_t valid_stride, int frame0, int frames, int n, int m, const float2 *__restrict__ tw,
                        float2 *__restrict__ zs) {
    __shared__ float2 lds[fft::Sizes<(1 << LOG2P)>::kLds];
    DeviceSync sync;
    fft::StagedPlan pl;
    pl.n = n;
    pl.m = m;
    pl.p = 1 << LOG2P;
    pl.log2p = LOG2P;
    fft::staged_first_sub<LOG2P>((int)threadIdx.x, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, frames, pl, in,
                                 station_pitch, frame0, valid, valid_stride, lds, sync, tw, tw + n, zs);
}

// Stage 2: grid (ceil((N/2+1)/256), frames, stations). Every form's words in
// the order of forms::spec_index (tile 16, 32 or N, the plain order): FORM 0
// complex64, FORM 1/2 bin k scaled by gain[st*N + k] and requantized.
constexpr unsigned kStagedBlock = 256;
template <int FORM>
__global__ void __launch_bounds__(kStagedBlock)
staged_r2c_kernel(const int32_t *__restrict__ valid, size_t valid_stride, int frame0, int frames, int n, int m,
                  int p, const float2 *__restrict__ zs, const float2 *__restrict__ wl,
                  const float *__restrict__ gain, void *__restrict__ out, size_t out_station_stride, int tile) {
    fft::StagedPlan pl;
    pl.n = n;
    pl.m = m;
    pl.p = p;
    const int f = (int)blockIdx.y, st = (int)blockIdx.z;
    const int frame = frame0 + f;
    const int fcap = (int)(out_station_stride / (size_t)n);
    auto at = [&](int k) { return forms::spec_index(st, frame, k, n, fcap, tile); };
    if constexpr (FORM == 0) {
        float2 *o = static_cast<float2 *>(out);
        fft::staged_r2c((int)(blockIdx.x * kStagedBlock + threadIdx.x), f, st, frames, pl, frame0, valid,
                        valid_stride, zs, wl, [&](int k, float2 x) { o[at(k)] = x; },
                        [&](int k) { o[at(k)] = make_float2(0.0f, 0.0f); });
    } else if constexpr (FORM == 1) {
        uint32_t *o = static_cast<uint32_t *>(out);
        const float *g = gain + (size_t)st * (size_t)n;
        fft::staged_r2c((int)(blockIdx.x * kStagedBlock + threadIdx.x), f, st, frames, pl, frame0, valid,
                        valid_stride, zs, wl, [&](int k, float2 x) { o[at(k)] = forms::pack_fp16(x, g[k]); },
                        [&](int k) { o[at(k)] = 0u; });
    } else {
        uint16_t *o = static_cast<uint16_t *>(out);
        const float *g = gain + (size_t)st * (size_t)n;
        fft::staged_r2c((int)(blockIdx.x * kStagedBlock + threadIdx.x), f, st, frames, pl, frame0, valid,
                        valid_stride, zs, wl, [&](int k, float2 x) { o[at(k)] = forms::pack_int8(x, g[k]); },
                        [&](int k) { o[at(k)] = (uint16_t)0; });
    }
}

// n_freq above 16384, the one-kernel form (fft_body.hpp staged_frame, used
// where staged_one_kernel holds): grid (m/2 + 1) * frames * stations, one
// dimension, mapped by fft::staged_unit; P/16 lanes, the P-point exchange buffer (139 KB at P = 16384). The FP16 (FORM 1)
// and int8 (FORM 2) forms' words, bin k scaled by gain[st*N + k] and
// requantized, in the order of forms::spec_index (tile 16, 32 or N).
template <int LOG2P, int FORM>
__global__ void __launch_bounds__((1 << LOG2P) / 16)
staged_kernel(const int8_t *__restrict__ in, size_t station_pitch, const int32_t *__restrict__ valid,
              size_t valid_stride, int frame0, int frames, int n_station, int n, int m,
              const float2 *__restrict__ tw, const float2 *__restrict__ wl, const float *__restrict__ gain,
              void *__restrict__ out, size_t out_station_stride, int tile) {
    __shared__ float2 lds[fft::Sizes<(1 << LOG2P)>::kLds];
    DeviceSync sync;
    fft::StagedPlan pl;
    pl.n = n;
    pl.m = m;
    pl.p = 1 << LOG2P;
    pl.log2p = LOG2P;
    unsigned qu, u;
    fft::staged_unit(blockIdx.x, (unsigned)(m / 2 + 1), (unsigned)frames * (unsigned)n_station, qu, u);
    const int q = (int)qu, f = (int)(u % (unsigned)frames), st = (int)(u / (unsigned)frames);
    const int frame = frame0 + f;
    const int fcap = (int)(out_station_stride / (size_t)n);
    auto at = [&](int k) { return forms::spec_index(st, frame, k, n, fcap, tile); };
    const int j = (int)threadIdx.x;
    if constexpr (FORM == 0) {
        float2 *o = static_cast<float2 *>(out);
        fft::staged_frame<LOG2P>(j, q, f, st, pl, in, station_pitch, frame0, valid, valid_stride, lds, sync, tw,
                                 tw + n, wl, [&](int k, float2 x) { o[at(k)] = x; },
                                 [&](int k) { o[at(k)] = make_float2(0.0f, 0.0f); });
    } else if constexpr (FORM == 1) {
        uint32_t *o = static_cast<uint32_t *>(out);
        const float *g = gain + (size_t)st * (size_t)n;
        fft::staged_frame<LOG2P>(j, q, f, st, pl, in, station_pitch, frame0, valid, valid_stride, lds, sync, tw,
                                 tw + n, wl, [&](int k, float2 x) { o[at(k)] = forms::pack_fp16(x, g[k]); },
                                 [&](int k) { o[at(k)] = 0u; });
    } else {
        uint16_t *o = static_cast<uint16_t *>(out);
        const float *g = gain + (size_t)st * (size_t)n;
        fft::staged_frame<LOG2P>(j, q, f, st, pl, in, station_pitch, frame0, valid, valid_stride, lds, sync, tw,
                                 tw + n, wl, [&](int k, float2 x) { o[at(k)] = forms::pack_int8(x, g[k]); },
                                 [&](int k) { o[at(k)] = (uint16_t)0; });
    }
}

// 4-bit packed samples -> int8, one 32-bit word (8 samples) per lane;
// grid (ceil(words/256), stations).
constexpr unsigned kUnpackBlock = 256;
__global__ void __launch_bounds__(kUnpackBlock)
unpack4_kernel(const uint32_t *__restrict__ in, size_t in_pitch, uint64_t *__restrict__ out, size_t out_pitch,
               size_t words) {
    fft::unpack4((size_t)blockIdx.x * kUnpackBlock + threadIdx.x, (int)blockIdx.y, in, in_pitch, out, out_pitch,
                 words);
}

// Validity bits -> int32 per (station, frame); grid (ceil(frames/256), stations).
constexpr unsigned kValidityBlock = 256;
__global__ void __launch_bounds__(kValidityBlock)
validity_kernel(const uint8_t *__restrict__ mask, size_t mask_pitch, int bit0, int frames,
                int32_t *__restrict__ valid, size_t valid_stride) {
    fft::expand_validity((int)(blockIdx.x * kValidityBlock + threadIdx.x), (int)blockIdx.y, frames,
                         mask, mask_pitch, bit0, valid, valid_stride);
}

// Twiddle tables in FP64: tw[m] = exp(-2 pi i m/n), wl[k] = exp(-2 pi i k/(2n)),
// and for p > 0 (n = M p, odd M) tw[n + m] = exp(-2 pi i m/p), m < p < n.
// The block size is a compile-time constant and the index is unsigned, so the
// kernel depends on no runtime-filled argument for its bounds.
constexpr unsigned kTwiddleBlock = 256;
__global__ void __launch_bounds__(kTwiddleBlock) twiddle_kernel(float2 *tw, float2 *wl, unsigned n, unsigned p) {
    const unsigned i = blockIdx.x * kTwiddleBlock + threadIdx.x;
    if (i >= n) return;
    double s, c;
    sincospi(-2.0 * (double)i / (double)n, &s, &c);
    tw[i] = make_float2((float)c, (float)s);
    sincospi(-(double)i / (double)n, &s, &c);
    wl[i] = make_float2((float)c, (float)s);
    if (i < p) {
        sincospi(-2.0 * (double)i / (double)p, &s, &c);
        tw[n + i] = make_float2((float)c, (float)s);
    }
}

// ---------------------------------------------------------------------------
// FP32 X-engine
// grid (ceil(nf/64), groups), block 64*waves; dynamic LDS [npad][frames_lds][64]
// ---------------------------------------------------------------------------
template <int NPAIR>
__global__ void __launch_bounds__(1024)
xengine_fp32_kernel(const float2 *__restrict__ y, size_t ys, int nf, int frames,
                    int n_station, const int2 *__restrict__ bpairs, int n_bpairs,
                    int frames_lds, float2 *__restrict__ acc, int nthreads) {
    extern __shared__ float2 sy[];
    DeviceSync sync;
    xeng::xengine_fp32_block<NPAIR>((int)threadIdx.x, nthreads, (int)blockIdx.x,
                                    (int)blockIdx.y, sy, sync, y, ys, nf, frames, n_station,
                                    bpairs, n_bpairs, frames_lds, acc);
}

// acc[i] *= s; grid-stride over `total` threads given explicitly (block 256).
constexpr unsigned kScaleBlock = 256;
__global__ void __launch_bounds__(kScaleBlock) scale_kernel(float2 *acc, size_t n, float s, size_t total) {
    for (size_t i = (size_t)blockIdx.x * kScaleBlock + threadIdx.x; i < n; i += total) {
        float2 v = acc[i];
        v.x *= s;
        v.y *= s;
        acc[i] = v;
    }
}

// ---------------------------------------------------------------------------
// MFMA X-engine (every form)
// grid (n_freq/CH), block 32*CH; dynamic LDS xmfma::lds_bytes(n_station, CH, KB).
// Up to 24 inputs (3 row tiles) with one staged item per lane, a CU holds 16
// waves of it (two 16-channel or one 32-channel workgroup), which keeps each
// lane to 128 registers; with two items per lane (NRT * KB > 4) the kernel may
// use more registers and holds fewer waves.
// ---------------------------------------------------------------------------
template <int FORM, int NRT, int CH, int KB>
__global__ void __launch_bounds__(xmfma::Tile<CH>::kThreads)
__attribute__((amdgpu_waves_per_eu(NRT <= 3 && NRT * KB <= 4 ? 4 : 1)))
xengine_mfma_kernel(const void *__restrict__ spectra, size_t ys, int nf, int spec_tile, int frames,
                    int n_station, int n_tiles, void *__restrict__ integ) {
    extern __shared__ uint4 lds_words[];
    DeviceSync sync;
    MfmaOps ops;
    xmfma::xengine_mfma_block<FORM, NRT, CH, KB>((int)threadIdx.x, (int)blockIdx.x, n_tiles,
                                                 reinterpret_cast<uint8_t *>(lds_words), sync, ops,
                                                 spectra, ys, nf, spec_tile, frames, n_station, integ);
}

template <int FORM, int NRT, int CH, int KB>
const void *xmfma_kernel_of() {
    if constexpr (xmfma::shape_built<NRT, CH, KB>())
        return reinterpret_cast<const void *>(&xengine_mfma_kernel<FORM, NRT, CH, KB>);
    else
        return nullptr;
}

template <int FORM, int CH, int KB>
const void *xmfma_kernel_of(int nrt) {
    switch (nrt) {
    case 1: return xmfma_kernel_of<FORM, 1, CH, KB>();
    case 2: return xmfma_kernel_of<FORM, 2, CH, KB>();
    case 3: return xmfma_kernel_of<FORM, 3, CH, KB>();
    case 4: return xmfma_kernel_of<FORM, 4, CH, KB>();
    case 5: return xmfma_kernel_of<FORM, 5, CH, KB>();
    case 6: return xmfma_kernel_of<FORM, 6, CH, KB>();
    default: return nullptr;
    }
}

// The X-engine kernel of (form, row tiles, channels, K blocks), or NULL.
const void *xmfma_kernel(int form, int nrt, int ch, int kb) {
#define CORR_XMFMA_OF(F, C, K) \
    if (form == F && ch == C && kb == K) return xmfma_kernel_of<F, C, K>(nrt);
#define CORR_XMFMA_FORM(F) \
    CORR_XMFMA_OF(F, 16, 1) CORR_XMFMA_OF(F, 16, 2) CORR_XMFMA_OF(F, 16, 4) \
    CORR_XMFMA_OF(F, 32, 1) CORR_XMFMA_OF(F, 32, 2) CORR_XMFMA_OF(F, 32, 4)
    CORR_XMFMA_FORM(0)
    CORR_XMFMA_FORM(1)
    CORR_XMFMA_FORM(2)
#undef CORR_XMFMA_FORM
#undef CORR_XMFMA_OF
    return nullptr;
}

// Calibration power sum and gains; grid (ceil(nf/256), stations), block 256.
constexpr unsigned kChanBlock = 256;
__global__ void __launch_bounds__(kChanBlock)
power_kernel(int nf, int frames, const float2 *__restrict__ spec, size_t spec_station_stride,
             const int32_t *__restrict__ valid, size_t valid_stride, float *__restrict__ power,
             int32_t *__restrict__ count) {
    fft::power_sum((int)(blockIdx.x * kChanBlock + threadIdx.x), (int)blockIdx.y, nf, frames, spec,
                   spec_station_stride, valid, valid_stride, power, count);
}

__global__ void __launch_bounds__(kChanBlock)
gains_kernel(int form, int nf, const float *__restrict__ power, const int32_t *__restrict__ count,
             float *__restrict__ gain) {
    const int c = (int)(blockIdx.x * kChanBlock + threadIdx.x), st = (int)blockIdx.y;
    if (c >= nf) return;
    const size_t at = (size_t)st * (size_t)nf + (size_t)c;
    gain[at] = fft::gain_from_power(form, power[at], count[st]);
}

// Fill n 32-bit words with v (grid-stride over `total` threads, block 256).
__global__ void __launch_bounds__(kChanBlock) fill_kernel(uint32_t *p, size_t n, uint32_t v, size_t total) {
    for (size_t i = (size_t)blockIdx.x * kChanBlock + threadIdx.x; i < n; i += total) p[i] = v;
}

// The autotuner's synthetic input: n words of fft::tuning_noise_word (grid-stride).
__global__ void __launch_bounds__(kChanBlock) noise_kernel(uint32_t *p, size_t n, size_t total) {
    for (size_t i = (size_t)blockIdx.x * kChanBlock + threadIdx.x; i < n; i += total) p[i] = fft::tuning_noise_word(i);
}

// Dump of one integration into complex64 [baseline][nf]; grid (ceil(nf/256), baselines).
//   FORM 0: acc * (1.0f/frames) in FP32 (the reference's row scale)
//   FORM 1: ((double)sum + (double)compensation) / (g_j g_k frames) in FP64
//   FORM 2: (double)int32 / (g_j g_k frames) in FP64
template <int FORM>
__global__ void __launch_bounds__(kChanBlock)
dump_kernel(const void *__restrict__ integ, int nf, const int2 *__restrict__ pairs,
            const float *__restrict__ gain, uint64_t frames, float inv_frames_f32,
            float2 *__restrict__ out) {
    const int c = (int)(blockIdx.x * kChanBlock + threadIdx.x), b = (int)blockIdx.y;
    if (c >= nf) return;
    const size_t at = (size_t)b * (size_t)nf + (size_t)c;
    if constexpr (FORM == 0) {
        const float2 v = static_cast<const float2 *>(integ)[at];
        out[at] = make_float2(v.x * inv_frames_f32, v.y * inv_frames_f32);
    } else {
        const int2 jk = pairs[b];
        const double den = (double)gain[(size_t)jk.x * nf + c] * (double)gain[(size_t)jk.y * nf + c] *
                           (double)frames;
        double re, im;
        if constexpr (FORM == 1) {
            const float4 v = static_cast<const float4 *>(integ)[at];
            re = ((double)v.x + (double)v.z) / den;
            im = ((double)v.y + (double)v.w) / den;
        } else {
            const int2 v = static_cast<const int2 *>(integ)[at];
            re = (double)v.x / den;
            im = (double)v.y / den;
        }
        out[at] = make_float2((float)re, (float)im);
    }
}

template <int LOG2N, int FORM>
hipError_t launch_fengine_form_t(const int8_t *in, size_t station_pitch, const int32_t *valid,
                                 size_t valid_stride, const float2 *tw, const float2 *wl,
                                 const float *gain, void *out, size_t out_station_stride,
                                 int tile_shift, int frames, int n_station, hipStream_t s) {
    const dim3 grid((unsigned)frames, (unsigned)n_station);
    const dim3 block((1u << LOG2N) / 16u);
    hipLaunchKernelGGL((fengine_form_kernel<LOG2N, FORM>), grid, block, 0, s, in, station_pitch,
                       valid, valid_stride, tw, wl, gain, out, out_station_stride, tile_shift);
    return hipGetLastError();
}

template <int M, int LOG2P, int FORM>
hipError_t launch_fengine_mixed_form_t(const int8_t *in, size_t station_pitch, const int32_t *valid,
                                       size_t valid_stride, const float2 *tw, const float2 *wl,
                                       const float *gain, void *out, size_t out_station_stride, int tile,
                                       int frames, int n_station, hipStream_t s) {
    const dim3 grid((unsigned)frames, (unsigned)n_station);
    const dim3 block((unsigned)fft::MixedSizes<M, LOG2P>::kThreads);
    hipLaunchKernelGGL((fengine_mixed_form_kernel<M, LOG2P, FORM>), grid, block, 0, s, in, station_pitch, valid,
                       valid_stride, tw, wl, gain, out, out_station_stride, tile);
    return hipGetLastError();
}

template <int M, int LOG2P>
hipError_t launch_fengine_mixed_fp32_t(const int8_t *in, size_t station_pitch, const int32_t *valid,
                                       size_t valid_stride, const float2 *tw, const float2 *wl, float2 *out,
                                       size_t out_station_stride, int tile, int frames, int n_station,
                                       hipStream_t s) {
    const dim3 grid((unsigned)frames, (unsigned)n_station);
    const dim3 block((unsigned)fft::MixedSizes<M, LOG2P>::kThreads);
    hipLaunchKernelGGL((fengine_mixed_fp32_kernel<M, LOG2P>), grid, block, 0, s, in, station_pitch, valid,
                       valid_stride, tw, wl, out, out_station_stride, tile);
    return hipGetLastError();
}

// tile: channels of the spectrum tile, 16, 32 or n_freq (the plain order).
template <int FORM>
hipError_t launch_fengine_form_f(int n_freq, const int8_t *in, size_t station_pitch,
                                 const int32_t *valid, size_t valid_stride, const float2 *tw,
                                 const float2 *wl, const float *gain, void *out,
                                 size_t out_station_stride, int tile, int frames,
                                 int n_station, hipStream_t s) {
    switch (n_freq) {
#define CORR_FFORM_CASE(L)                                                                       \
    case 1 << L:                                                                                 \
        return launch_fengine_form_t<L, FORM>(in, station_pitch, valid, valid_stride, tw, wl,   \
                                              gain, out, out_station_stride,                    \
                                              tile == 16 ? 4 : tile == 32 ? 5 : L, frames,      \
                                              n_station, s);
        CORR_FFORM_CASE(9)
        CORR_FFORM_CASE(10)
        CORR_FFORM_CASE(11)
        CORR_FFORM_CASE(12)
        CORR_FFORM_CASE(13)
        CORR_FFORM_CASE(14)
#undef CORR_FFORM_CASE
#define CORR_FMIXED_CASE(M, L)                                                                            \
    case M << L:                                                                                          \
        return launch_fengine_mixed_form_t<M, L, FORM>(in, station_pitch, valid, valid_stride, tw, wl, gain, \
                                                       out, out_station_stride, tile, frames, n_station, s);
        CORR_FMIXED_CASE(3, 11)
        CORR_FMIXED_CASE(5, 11)
        CORR_FMIXED_CASE(3, 12)
        CORR_FMIXED_CASE(7, 11)
#undef CORR_FMIXED_CASE
    default:
        return hipErrorInvalidValue;
    }
}

template <int LOG2N>
hipError_t launch_fengine_fp32_t(const int8_t *in, size_t station_pitch, const int32_t *valid,
                                 size_t valid_stride, const float2 *tw, const float2 *wl,
                                 float2 *out, size_t out_station_stride, int tile_shift, int frames,
                                 int n_station, hipStream_t s) {
    const dim3 grid((unsigned)frames, (unsigned)n_station);
    const dim3 block((1u << LOG2N) / 16u);
    hipLaunchKernelGGL(fengine_fp32_kernel<LOG2N>, grid, block, 0, s, in, station_pitch, valid,
                       valid_stride, tw, wl, out, out_station_stride, tile_shift);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_fengine_fp32(int n_freq, const int8_t *in, size_t station_pitch,
                               const int32_t *valid, size_t valid_stride, const float2 *tw,
                               const float2 *wl, float2 *out, size_t out_station_stride, int tile,
                               int frames, int n_station, hipStream_t s) {
    if (frames <= 0 || n_station <= 0) return hipSuccess;
    // a tile of n_freq channels is the plain order [station][frame][channel]
    if (tile != 16 && tile != 32 && tile != n_freq) return hipErrorInvalidValue;
    switch (n_freq) {
#define CORR_FENGINE_CASE(L)                                                                      \
    case 1 << L:                                                                                  \
        return launch_fengine_fp32_t<L>(in, station_pitch, valid, valid_stride, tw, wl, out, \
                                        out_station_stride, tile == 16 ? 4 : tile == 32 ? 5 : L,  \
                                        frames, n_station, s);
        CORR_FENGINE_CASE(9)
        CORR_FENGINE_CASE(10)
        CORR_FENGINE_CASE(11)
        CORR_FENGINE_CASE(12)
        CORR_FENGINE_CASE(13)
        CORR_FENGINE_CASE(14)
#undef CORR_FENGINE_CASE
#define CORR_FMIXED_CASE(M, L)                                                                                \
    case M << L:                                                                                              \
        return launch_fengine_mixed_fp32_t<M, L>(in, station_pitch, valid, valid_stride, tw, wl, out, \
                                                 out_station_stride, tile, frames, n_station, s);
        CORR_FMIXED_CASE(3, 11)
        CORR_FMIXED_CASE(5, 11)
        CORR_FMIXED_CASE(3, 12)
        CORR_FMIXED_CASE(7, 11)
#undef CORR_FMIXED_CASE
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_fengine_staged(int form, int n_freq, const int8_t *in, size_t station_pitch, const int32_t *valid,
                                 size_t valid_stride, int frame0, int frames, int n_station, const float2 *tw,
                                 const float2 *wl, const float *gain, void *out, size_t out_station_stride, int tile,
                                 float2 *zs, hipStream_t s) {
    if (frames <= 0 || n_station <= 0) return hipSuccess;
    const fft::StagedPlan pl = fft::staged_plan(n_freq);
    if (pl.m == 0 || form < 0 || form > 2 || (tile != 16 && tile != 32 && tile != n_freq) || frames > 65535 ||
        n_station > 65535)
        return hipErrorInvalidValue;
    if (fft::staged_one_kernel(form, pl)) {
        const size_t groups = (size_t)(pl.m / 2 + 1) * (size_t)frames * (size_t)n_station;
        if (groups > 0x7fffffffu) return hipErrorInvalidValue;
        const dim3 grid((unsigned)groups), block((unsigned)(pl.p / 16));
        switch (pl.log2p * 4 + form) {
#define CORR_STAGED1_CASE(L, F)                                                                                      \
    case L * 4 + F:                                                                                                   \
        hipLaunchKernelGGL((staged_kernel<L, F>), grid, block, 0, s, in, station_pitch, valid, valid_stride, frame0,   \
                           frames, n_station, pl.n, pl.m, tw, wl, gain, out, out_station_stride, tile);              \
        break;
            CORR_STAGED1_CASE(11, 0)
            CORR_STAGED1_CASE(11, 1)
            CORR_STAGED1_CASE(11, 2)
            CORR_STAGED1_CASE(12, 0)
            CORR_STAGED1_CASE(12, 1)
            CORR_STAGED1_CASE(12, 2)
            CORR_STAGED1_CASE(13, 0)
            CORR_STAGED1_CASE(13, 1)
            CORR_STAGED1_CASE(13, 2)
            CORR_STAGED1_CASE(14, 0)
            CORR_STAGED1_CASE(14, 1)
            CORR_STAGED1_CASE(14, 2)
#undef CORR_STAGED1_CASE
        default:
            return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    if ((size_t)frames * (size_t)n_station > 65535u) return hipErrorInvalidValue;  // the two-kernel grids
    const dim3 sub_grid((unsigned)pl.m, (unsigned)frames, (unsigned)n_station);
    const dim3 sub_block((unsigned)(pl.p / 16));
    hipError_t e;
    switch (pl.log2p) {
#define CORR_STAGED_CASE(L)                                                                                          \
    case L:                                                                                                           \
        hipLaunchKernelGGL(staged_first_sub_kernel<L>, sub_grid, sub_block, 0, s, in, station_pitch, valid,            \
                           valid_stride, frame0, frames, pl.n, pl.m, tw, zs);                                        \
        break;
        CORR_STAGED_CASE(11)
        CORR_STAGED_CASE(12)
        CORR_STAGED_CASE(13)
        CORR_STAGED_CASE(14)
#undef CORR_STAGED_CASE
    default:
        return hipErrorInvalidValue;
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const dim3 r2c_grid((unsigned)((n_freq / 2 + 1 + kStagedBlock - 1) / kStagedBlock), (unsigned)frames,
                        (unsigned)n_station);
    if (form == 0)
        hipLaunchKernelGGL(staged_r2c_kernel<0>, r2c_grid, dim3(kStagedBlock), 0, s, valid, valid_stride, frame0,
                           frames, pl.n, pl.m, pl.p, zs, wl, gain, out, out_station_stride, tile);
    else if (form == 1)
        hipLaunchKernelGGL(staged_r2c_kernel<1>, r2c_grid, dim3(kStagedBlock), 0, s, valid, valid_stride, frame0,
                           frames, pl.n, pl.m, pl.p, zs, wl, gain, out, out_station_stride, tile);
    else
        hipLaunchKernelGGL(staged_r2c_kernel<2>, r2c_grid, dim3(kStagedBlock), 0, s, valid, valid_stride, frame0,
                           frames, pl.n, pl.m, pl.p, zs, wl, gain, out, out_station_stride, tile);
    return hipGetLastError();
}

hipError_t launch_unpack4(const int8_t *in, size_t in_pitch, int8_t *out, size_t out_pitch, int frames,
                          int n_freq, int n_station, hipStream_t s) {
    if (frames <= 0 || n_station <= 0) return hipSuccess;
    // a packed frame is n_freq bytes, n_freq/4 words; pitches are whole words
    if (n_freq % 4 != 0 || in_pitch % 4 != 0 || out_pitch % 8 != 0 ||
        in_pitch < (size_t)frames * (size_t)n_freq || out_pitch < (size_t)frames * 2u * (size_t)n_freq)
        return hipErrorInvalidValue;
    const size_t words = (size_t)frames * (size_t)n_freq / 4u;
    const dim3 grid((unsigned)((words + kUnpackBlock - 1) / kUnpackBlock), (unsigned)n_station);
    hipLaunchKernelGGL(unpack4_kernel, grid, dim3(kUnpackBlock), 0, s, reinterpret_cast<const uint32_t *>(in),
                       in_pitch / 4u, reinterpret_cast<uint64_t *>(out), out_pitch / 8u, words);
    return hipGetLastError();
}

hipError_t launch_twiddles(float2 *tw, float2 *wl, int n_freq, hipStream_t s) {
    if (n_freq <= 0) return hipErrorInvalidValue;
    const unsigned p = (unsigned)(fft_twiddle_elems(n_freq) - n_freq);
    hipLaunchKernelGGL(twiddle_kernel, dim3((unsigned)((n_freq + kTwiddleBlock - 1) / kTwiddleBlock)),
                       dim3(kTwiddleBlock), 0, s, tw, wl, (unsigned)n_freq, p);
    return hipGetLastError();
}

hipError_t launch_validity(const uint8_t *mask, size_t mask_pitch, int bit0, int frames,
                           int n_station, int32_t *valid, size_t valid_stride, hipStream_t s) {
    if (frames <= 0 || n_station <= 0) return hipSuccess;
    const dim3 grid((unsigned)((frames + kValidityBlock - 1) / kValidityBlock), (unsigned)n_station);
    hipLaunchKernelGGL(validity_kernel, grid, dim3(kValidityBlock), 0, s, mask, mask_pitch, bit0,
                       frames, valid, valid_stride);
    return hipGetLastError();
}

hipError_t launch_xengine_fp32(const float2 *y, size_t y_station_stride, int n_freq, int frames,
                               int n_station, const int2 *bpairs, int n_bpairs, int npair,
                               int waves, int groups, int frames_lds, float2 *acc,
                               hipStream_t s) {
    if (frames <= 0) return hipSuccess;
    const size_t shm = xeng::lds_bytes(n_station, frames_lds);
    const dim3 grid((unsigned)((n_freq + 63) / 64), (unsigned)groups);
    const dim3 block((unsigned)(64 * waves));
    switch (npair) {
#define CORR_XENGINE_CASE(P)                                                                  \
    case P:                                                                                   \
        hipLaunchKernelGGL(xengine_fp32_kernel<P>, grid, block, shm, s, y, y_station_stride, \
                           n_freq, frames, n_station, bpairs, n_bpairs, frames_lds, acc,      \
                           64 * waves);                                                       \
        return hipGetLastError();
        CORR_XENGINE_CASE(1)
        CORR_XENGINE_CASE(2)
#undef CORR_XENGINE_CASE
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_fengine_form(int form, int n_freq, const int8_t *in, size_t station_pitch,
                               const int32_t *valid, size_t valid_stride, const float2 *tw,
                               const float2 *wl, const float *gain, void *out,
                               size_t out_station_stride, int tile, int frames, int n_station,
                               hipStream_t s) {
    if (frames <= 0 || n_station <= 0) return hipSuccess;
    // a tile of n_freq channels is the plain order [station][frame][channel]
    if (tile != 16 && tile != 32 && tile != n_freq) return hipErrorInvalidValue;
    if (form == 1)
        return launch_fengine_form_f<1>(n_freq, in, station_pitch, valid, valid_stride, tw, wl, gain,
                                        out, out_station_stride, tile, frames, n_station, s);
    if (form == 2)
        return launch_fengine_form_f<2>(n_freq, in, station_pitch, valid, valid_stride, tw, wl, gain,
                                        out, out_station_stride, tile, frames, n_station, s);
    return hipErrorInvalidValue;
}

hipError_t launch_power(int nf, int frames, int n_station, const float2 *spec,
                        size_t spec_station_stride, const int32_t *valid, size_t valid_stride,
                        float *power, int32_t *count, hipStream_t s) {
    if (frames <= 0 || n_station <= 0 || nf <= 0) return hipSuccess;
    const dim3 grid((unsigned)((nf + kChanBlock - 1) / kChanBlock), (unsigned)n_station);
    hipLaunchKernelGGL(power_kernel, grid, dim3(kChanBlock), 0, s, nf, frames, spec,
                       spec_station_stride, valid, valid_stride, power, count);
    return hipGetLastError();
}

hipError_t launch_gains(int form, int nf, int n_station, const float *power, const int32_t *count,
                        float *gain, hipStream_t s) {
    if (n_station <= 0 || nf <= 0) return hipSuccess;
    const dim3 grid((unsigned)((nf + kChanBlock - 1) / kChanBlock), (unsigned)n_station);
    hipLaunchKernelGGL(gains_kernel, grid, dim3(kChanBlock), 0, s, form, nf, power, count, gain);
    return hipGetLastError();
}

hipError_t launch_fill(void *p, size_t n_words, uint32_t v, hipStream_t s) {
    if (n_words == 0) return hipSuccess;
    size_t blocks = (n_words + kChanBlock - 1) / kChanBlock;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(kChanBlock), 0, s,
                       static_cast<uint32_t *>(p), n_words, v, blocks * kChanBlock);
    return hipGetLastError();
}

hipError_t launch_noise(void *p, size_t n_words, hipStream_t s) {
    if (n_words == 0) return hipSuccess;
    size_t blocks = (n_words + kChanBlock - 1) / kChanBlock;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(noise_kernel, dim3((unsigned)blocks), dim3(kChanBlock), 0, s, static_cast<uint32_t *>(p), n_words,
                       blocks * kChanBlock);
    return hipGetLastError();
}

hipError_t xengine_mfma_prepare(int form, int n_station, int ch, int kb) {
    if (!xmfma::tile_fits(n_station, ch, kb)) return hipErrorInvalidValue;
    const void *fn = xmfma_kernel(form, xmfma::row_tiles(n_station), ch, kb);
    if (fn == nullptr) return hipErrorInvalidValue;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)xmfma::lds_bytes(n_station, ch, kb));
}

hipError_t launch_xengine_mfma(int form, int n_station, int ch, int kb, const void *spectra, size_t ys, int nf,
                               int spec_tile, int frames, void *integ, hipStream_t s) {
    if (frames <= 0) return hipSuccess;
    if (!xmfma::tile_fits(n_station, ch, kb) || nf % ch != 0) return hipErrorInvalidValue;
    if (spec_tile != ch && spec_tile != nf) return hipErrorInvalidValue;
    const void *fn = xmfma_kernel(form, xmfma::row_tiles(n_station), ch, kb);
    if (fn == nullptr) return hipErrorInvalidValue;
    int n_tiles = nf / ch;
    const size_t shm = xmfma::lds_bytes(n_station, ch, kb);
    // the kernel's arguments in order, each by address (hipLaunchKernel)
    void *args[] = {&spectra, &ys, &nf, &spec_tile, &frames, &n_station, &n_tiles, &integ};
    const hipError_t e = hipLaunchKernel(fn, dim3((unsigned)n_tiles), dim3((unsigned)(32 * ch)), args, shm, s);
    return e != hipSuccess ? e : hipGetLastError();
}

hipError_t launch_dump(int form, const void *integ, int n_baseline, int nf, const int2 *pairs,
                       const float *gain, uint64_t frames, float2 *out, hipStream_t s) {
    if (n_baseline <= 0 || nf <= 0) return hipSuccess;
    const dim3 grid((unsigned)((nf + kChanBlock - 1) / kChanBlock), (unsigned)n_baseline);
    const float inv = 1.0f / (float)frames;
    switch (form) {
    case 0:
        hipLaunchKernelGGL(dump_kernel<0>, grid, dim3(kChanBlock), 0, s, integ, nf, pairs, gain,
                           frames, inv, out);
        break;
    case 1:
        hipLaunchKernelGGL(dump_kernel<1>, grid, dim3(kChanBlock), 0, s, integ, nf, pairs, gain,
                           frames, inv, out);
        break;
    case 2:
        hipLaunchKernelGGL(dump_kernel<2>, grid, dim3(kChanBlock), 0, s, integ, nf, pairs, gain,
                           frames, inv, out);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_scale(float2 *acc, size_t n, float s, hipStream_t st) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + kScaleBlock - 1) / kScaleBlock;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(scale_kernel, dim3((unsigned)blocks), dim3(kScaleBlock), 0, st, acc, n, s,
                       blocks * kScaleBlock);
    return hipGetLastError();
}

}  // namespace corr
