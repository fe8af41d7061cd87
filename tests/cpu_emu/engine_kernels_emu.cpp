// The engine's kernel launchers (declared in csrc/engine.hpp) for the CPU build:
// each launch runs the very kernel bodies of csrc/fft_body.hpp and
// csrc/xengine_body.hpp over the launch's grid, one host thread per lane of a
// workgroup and a barrier for the workgroup barrier, with the launch geometry
// (grid, block, LDS bytes) the device launchers use.
#include <hip/hip_runtime.h>

#include <cmath>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "fft_body.hpp"
#include "xengine_body.hpp"

namespace corr {
namespace {

// Run body(tid, bx, by, lds, sync) for every workgroup of a (gx, gy) grid.
template <class Body>
void run_grid(int gx, int gy, int nthreads, size_t lds_bytes, Body body) {
    std::barrier<> bar(nthreads);
    std::vector<std::thread> th;
    char *lds = new char[lds_bytes ? lds_bytes : 1];
    for (int t = 0; t < nthreads; ++t)
        th.emplace_back([&, t] {
            HostSync sync{&bar};
            for (int by = 0; by < gy; ++by)
                for (int bx = 0; bx < gx; ++bx) {
                    body(t, bx, by, lds, sync);
                    sync();  // the next workgroup reuses the LDS
                }
        });
    for (auto &x : th) x.join();
    delete[] lds;
}

template <int LOG2N>
hipError_t fengine_t(const int8_t *in, size_t station_pitch, const int32_t *valid, size_t valid_stride,
                     const float2 *tw, const float2 *wl, float2 *out, size_t out_station_stride,
                     int frames, int n_station) {
    constexpr int N = 1 << LOG2N;
    run_grid(frames, n_station, N / 16, fft::Sizes<N>::kLds * sizeof(float2),
             [=](int j, int bx, int by, char *lds, HostSync &sync) {
                 fft::fengine_fp32_block<LOG2N>(j, bx, by, reinterpret_cast<float2 *>(lds), sync, in,
                                                station_pitch, valid, valid_stride, tw, wl, out,
                                                out_station_stride);
             });
    return hipSuccess;
}

}  // namespace

hipError_t launch_twiddles(float2 *tw, float2 *wl, int n_freq, hipStream_t) {
    if (n_freq <= 0) return hipErrorInvalidValue;
    for (int i = 0; i < n_freq; ++i) {
        const double a = -2.0 * M_PI * i / n_freq, b = -M_PI * i / n_freq;
        tw[i] = make_float2((float)std::cos(a), (float)std::sin(a));
        wl[i] = make_float2((float)std::cos(b), (float)std::sin(b));
    }
    return hipSuccess;
}

hipError_t launch_validity(const uint8_t *mask, size_t mask_pitch, int bit0, int frames,
                           int n_station, int32_t *valid, size_t valid_stride, hipStream_t) {
    if (frames <= 0 || n_station <= 0) return hipSuccess;
    const int gx = (frames + 255) / 256;  // the device launcher's grid, 256 lanes
    for (int st = 0; st < n_station; ++st)
        for (int lane = 0; lane < gx * 256; ++lane)
            fft::expand_validity(lane, st, frames, mask, mask_pitch, bit0, valid, valid_stride);
    return hipSuccess;
}

hipError_t launch_fengine_fp32(int log2_nf, const int8_t *in, size_t station_pitch,
                               const int32_t *valid, size_t valid_stride, const float2 *tw,
                               const float2 *wl, float2 *out, size_t out_station_stride,
                               int frames, int n_station, hipStream_t) {
    if (frames <= 0 || n_station <= 0) return hipSuccess;
    switch (log2_nf) {
#define E(L)  \
    case L: \
        return fengine_t<L>(in, station_pitch, valid, valid_stride, tw, wl, out, out_station_stride, frames, n_station);
        E(9) E(10) E(11) E(12) E(13) E(14)
#undef E
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_xengine_fp32(const float2 *y, size_t y_station_stride, int n_freq, int frames,
                               int n_station, const int2 *bpairs, int n_bpairs, int npair,
                               int waves, int groups, int frames_lds, float2 *acc, hipStream_t) {
    if (frames <= 0) return hipSuccess;
    const int nthreads = 64 * waves;
    const size_t shm = xeng::lds_bytes(n_station, frames_lds);
    auto body = [=](int t, int bx, int by, char *lds, HostSync &sync) {
        float2 *sy = reinterpret_cast<float2 *>(lds);
        if (npair == 1)
            xeng::xengine_fp32_block<1>(t, nthreads, bx, by, sy, sync, y, y_station_stride, n_freq,
                                        frames, n_station, bpairs, n_bpairs, frames_lds, acc);
        else
            xeng::xengine_fp32_block<2>(t, nthreads, bx, by, sy, sync, y, y_station_stride, n_freq,
                                        frames, n_station, bpairs, n_bpairs, frames_lds, acc);
    };
    if (npair != 1 && npair != 2) return hipErrorInvalidValue;
    run_grid((n_freq + 63) / 64, groups, nthreads, shm, body);
    return hipSuccess;
}

hipError_t launch_scale(float2 *acc, size_t n, float s, hipStream_t) {
    for (size_t i = 0; i < n; ++i) {
        acc[i].x *= s;
        acc[i].y *= s;
    }
    return hipSuccess;
}

}  // namespace corr
