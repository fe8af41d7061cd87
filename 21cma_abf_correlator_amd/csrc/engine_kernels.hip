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

namespace corr {
namespace {

struct DeviceSync {
    __device__ void operator()() const { __syncthreads(); }
};

// ---------------------------------------------------------------------------
// F-engine, complex64 output
// grid (frames, stations), block N/16 threads, N = n_freq = 2^LOG2N
// ---------------------------------------------------------------------------
template <int LOG2N>
__global__ void __launch_bounds__((1 << LOG2N) / 16)
fengine_fp32_kernel(const int8_t *__restrict__ in, size_t station_pitch,
                    const int32_t *__restrict__ valid, size_t valid_stride,
                    const float2 *__restrict__ tw, const float2 *__restrict__ wl,
                    float2 *__restrict__ out, size_t out_station_stride) {
    __shared__ float2 lds[fft::Sizes<(1 << LOG2N)>::kLds];
    DeviceSync sync;
    fft::fengine_fp32_block<LOG2N>((int)threadIdx.x, (int)blockIdx.x, (int)blockIdx.y, lds, sync, in,
                                   station_pitch, valid, valid_stride, tw, wl, out,
                                   out_station_stride);
}

// Validity bits -> int32 per (station, frame); grid (ceil(frames/256), stations).
constexpr unsigned kValidityBlock = 256;
__global__ void __launch_bounds__(kValidityBlock)
validity_kernel(const uint8_t *__restrict__ mask, size_t mask_pitch, int bit0, int frames,
                int32_t *__restrict__ valid, size_t valid_stride) {
    fft::expand_validity((int)(blockIdx.x * kValidityBlock + threadIdx.x), (int)blockIdx.y, frames,
                         mask, mask_pitch, bit0, valid, valid_stride);
}

// Twiddle tables in FP64: tw[m] = exp(-2 pi i m/n), wl[k] = exp(-2 pi i k/(2n)).
// The block size is a compile-time constant and the index is unsigned, so the
// kernel depends on no runtime-filled argument for its bounds.
constexpr unsigned kTwiddleBlock = 256;
__global__ void __launch_bounds__(kTwiddleBlock) twiddle_kernel(float2 *tw, float2 *wl, unsigned n) {
    const unsigned i = blockIdx.x * kTwiddleBlock + threadIdx.x;
    if (i >= n) return;
    double s, c;
    sincospi(-2.0 * (double)i / (double)n, &s, &c);
    tw[i] = make_float2((float)c, (float)s);
    sincospi(-(double)i / (double)n, &s, &c);
    wl[i] = make_float2((float)c, (float)s);
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

template <int LOG2N>
hipError_t launch_fengine_fp32_t(const int8_t *in, size_t station_pitch, const int32_t *valid,
                                 size_t valid_stride, const float2 *tw, const float2 *wl,
                                 float2 *out, size_t out_station_stride, int frames,
                                 int n_station, hipStream_t s) {
    const dim3 grid((unsigned)frames, (unsigned)n_station);
    const dim3 block((1u << LOG2N) / 16u);
    hipLaunchKernelGGL(fengine_fp32_kernel<LOG2N>, grid, block, 0, s, in, station_pitch, valid,
                       valid_stride, tw, wl, out, out_station_stride);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_twiddles(float2 *tw, float2 *wl, int n_freq, hipStream_t s) {
    if (n_freq <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(twiddle_kernel, dim3((unsigned)((n_freq + kTwiddleBlock - 1) / kTwiddleBlock)),
                       dim3(kTwiddleBlock), 0, s, tw, wl, (unsigned)n_freq);
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

hipError_t launch_fengine_fp32(int log2_nf, const int8_t *in, size_t station_pitch,
                               const int32_t *valid, size_t valid_stride, const float2 *tw,
                               const float2 *wl, float2 *out, size_t out_station_stride,
                               int frames, int n_station, hipStream_t s) {
    if (frames <= 0 || n_station <= 0) return hipSuccess;
    switch (log2_nf) {
#define CORR_FENGINE_CASE(L)                                                              \
    case L:                                                                               \
        return launch_fengine_fp32_t<L>(in, station_pitch, valid, valid_stride, tw, wl, out, \
                                        out_station_stride, frames, n_station, s);
        CORR_FENGINE_CASE(9)
        CORR_FENGINE_CASE(10)
        CORR_FENGINE_CASE(11)
        CORR_FENGINE_CASE(12)
        CORR_FENGINE_CASE(13)
        CORR_FENGINE_CASE(14)
#undef CORR_FENGINE_CASE
    default:
        return hipErrorInvalidValue;
    }
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

hipError_t launch_scale(float2 *acc, size_t n, float s, hipStream_t st) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + kScaleBlock - 1) / kScaleBlock;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(scale_kernel, dim3((unsigned)blocks), dim3(kScaleBlock), 0, st, acc, n, s,
                       blocks * kScaleBlock);
    return hipGetLastError();
}

}  // namespace corr
