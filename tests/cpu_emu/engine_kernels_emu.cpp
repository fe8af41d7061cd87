// The engine's kernel launchers (declared in csrc/engine.hpp) for the CPU build:
// each launch runs the very kernel bodies of csrc/fft_body.hpp,
// csrc/xengine_body.hpp and csrc/xmfma_body.hpp over the launch's grid, one host thread per lane of a
// workgroup and a barrier for the workgroup barrier, with the launch geometry
// (grid, block, LDS bytes) the device launchers use.
#include <hip/hip_runtime.h>

#include <cmath>
#include <thread>
#include <vector>

#include "emu_mfma.hpp"
#include "engine.hpp"
#include "fft_body.hpp"
#include "xengine_body.hpp"
#include "xmfma_body.hpp"

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
    run_grid(frames, n_station, N / 16, fft::fengine_lds_elems<LOG2N>() * sizeof(float2),
             [=](int j, int bx, int by, char *lds, HostSync &sync) {
                 fft::fengine_fp32_block<LOG2N>(j, bx, by, reinterpret_cast<float2 *>(lds), sync, in,
                                                station_pitch, valid, valid_stride, tw, wl, out,
                                                out_station_stride);
             });
    return hipSuccess;
}

template <int LOG2N, int FORM>
hipError_t fengine_form_t(const int8_t *in, size_t station_pitch, const int32_t *valid, size_t valid_stride,
                          const float2 *tw, const float2 *wl, const float *gain, void *out,
                          size_t out_station_stride, int tile_shift, int frames, int n_station) {
    constexpr int N = 1 << LOG2N;
    run_grid(frames, n_station, N / 16, fft::fengine_lds_elems<LOG2N>() * sizeof(float2),
             [=](int j, int bx, int by, char *lds, HostSync &sync) {
                 fft::fengine_form_block<LOG2N, FORM>(j, bx, by, reinterpret_cast<float2 *>(lds), sync, in,
                                                      station_pitch, valid, valid_stride, tw, wl, gain, out,
                                                      out_station_stride, tile_shift);
             });
    return hipSuccess;
}

template <int M, int LOG2P>
hipError_t fengine_mixed_t(const int8_t *in, size_t station_pitch, const int32_t *valid, size_t valid_stride,
                           const float2 *tw, const float2 *wl, float2 *out, size_t out_station_stride, int frames,
                           int n_station) {
    using S = fft::MixedSizes<M, LOG2P>;
    run_grid(frames, n_station, S::kThreads, S::kLds * sizeof(float2),
             [=](int j, int bx, int by, char *lds, HostSync &sync) {
                 fft::fengine_mixed_fp32_block<M, LOG2P>(j, bx, by, reinterpret_cast<float2 *>(lds), sync, in,
                                                         station_pitch, valid, valid_stride, tw, wl, out,
                                                         out_station_stride);
             });
    return hipSuccess;
}

template <int M, int LOG2P, int FORM>
hipError_t fengine_mixed_form_t(const int8_t *in, size_t station_pitch, const int32_t *valid, size_t valid_stride,
                                const float2 *tw, const float2 *wl, const float *gain, void *out,
                                size_t out_station_stride, int tile, int frames, int n_station) {
    using S = fft::MixedSizes<M, LOG2P>;
    run_grid(frames, n_station, S::kThreads, S::kLds * sizeof(float2),
             [=](int j, int bx, int by, char *lds, HostSync &sync) {
                 fft::fengine_mixed_form_block<M, LOG2P, FORM>(j, bx, by, reinterpret_cast<float2 *>(lds), sync,
                                                               in, station_pitch, valid, valid_stride, tw, wl, gain,
                                                               out, out_station_stride, tile);
             });
    return hipSuccess;
}

template <int FORM, int NRT, int CH, int KB>
void xmfma_t(int n_station, const void *spectra, size_t ys, int nf, int spec_tile, int frames, void *integ) {
    if constexpr (xmfma::shape_built<NRT, CH, KB>()) {  // the others are never admitted by tile_fits
        const int n_tiles = nf / CH;
        EmuMfmaShared *sh = new EmuMfmaShared;
        run_grid(n_tiles, 1, xmfma::Tile<CH>::kThreads, xmfma::lds_bytes(n_station, CH, KB),
                 [=](int t, int bx, int, char *lds, HostSync &sync) {
                     EmuMfmaOps ops{t, sh, &sync};
                     xmfma::xengine_mfma_block<FORM, NRT, CH, KB>(t, bx, n_tiles, reinterpret_cast<uint8_t *>(lds),
                                                                  sync, ops, spectra, ys, nf, spec_tile, frames,
                                                                  n_station, integ);
                 });
        delete sh;
    }
}

// The row-tile counts of a (channels, K blocks) shape, as the GPU build has
// them (tile_fits has been checked by the caller).
template <int FORM, int CH, int KB>
hipError_t xmfma_nrt(int nrt, int n_station, const void *spectra, size_t ys, int nf, int tl, int frames,
                     void *integ) {
    switch (nrt) {
    case 1: xmfma_t<FORM, 1, CH, KB>(n_station, spectra, ys, nf, tl, frames, integ); return hipSuccess;
    case 2: xmfma_t<FORM, 2, CH, KB>(n_station, spectra, ys, nf, tl, frames, integ); return hipSuccess;
    case 3: xmfma_t<FORM, 3, CH, KB>(n_station, spectra, ys, nf, tl, frames, integ); return hipSuccess;
    case 4: xmfma_t<FORM, 4, CH, KB>(n_station, spectra, ys, nf, tl, frames, integ); return hipSuccess;
    case 5: xmfma_t<FORM, 5, CH, KB>(n_station, spectra, ys, nf, tl, frames, integ); return hipSuccess;
    case 6: xmfma_t<FORM, 6, CH, KB>(n_station, spectra, ys, nf, tl, frames, integ); return hipSuccess;
    default: return hipErrorInvalidValue;
    }
}

template <int FORM>
hipError_t xmfma_f(int n_station, int ch, int kb, const void *spectra, size_t ys, int nf, int tl, int frames,
                   void *integ) {
    const int nrt = xmfma::row_tiles(n_station);
#define E(C, K) \
    if (ch == C && kb == K) return xmfma_nrt<FORM, C, K>(nrt, n_station, spectra, ys, nf, tl, frames, integ);
    E(16, 1) E(16, 2) E(16, 4) E(32, 1) E(32, 2) E(32, 4)
#undef E
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_twiddles(float2 *tw, float2 *wl, int n_freq, hipStream_t) {
    if (n_freq <= 0) return hipErrorInvalidValue;
    for (int i = 0; i < n_freq; ++i) {
        const double a = -2.0 * M_PI * i / n_freq, b = -M_PI * i / n_freq;
        tw[i] = make_float2((float)std::cos(a), (float)std::sin(a));
        wl[i] = make_float2((float)std::cos(b), (float)std::sin(b));
    }
    const int p = fft_twiddle_elems(n_freq) - n_freq;  // the sub-transforms' table
    for (int i = 0; i < p; ++i) {
        const double a = -2.0 * M_PI * i / p;
        tw[n_freq + i] = make_float2((float)std::cos(a), (float)std::sin(a));
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

hipError_t launch_fengine_fp32(int n_freq, const int8_t *in, size_t station_pitch,
                               const int32_t *valid, size_t valid_stride, const float2 *tw,
                               const float2 *wl, float2 *out, size_t out_station_stride, int tile,
                               int frames, int n_station, hipStream_t) {
    if (frames <= 0 || n_station <= 0) return hipSuccess;
    if (tile != 16 && tile != 32 && tile != n_freq) return hipErrorInvalidValue;
    switch (n_freq) {
#define E(L)                                                                                                    \
    case 1 << L:                                                                                                \
        if (tile == n_freq)                                                                                     \
            return fengine_t<L>(in, station_pitch, valid, valid_stride, tw, wl, out, out_station_stride, frames, \
                                n_station);                                                                     \
        return fengine_form_t<L, 0>(in, station_pitch, valid, valid_stride, tw, wl, nullptr, out,              \
                                    out_station_stride, tile == 16 ? 4 : 5, frames, n_station);
        E(9) E(10) E(11) E(12) E(13) E(14)
#undef E
#define E(M, L)                                                                                                   \
    case M << L:                                                                                                  \
        if (tile == n_freq)                                                                                       \
            return fengine_mixed_t<M, L>(in, station_pitch, valid, valid_stride, tw, wl, out, out_station_stride, \
                                         frames, n_station);                                                      \
        return fengine_mixed_form_t<M, L, 0>(in, station_pitch, valid, valid_stride, tw, wl, nullptr, out,       \
                                             out_station_stride, tile, frames, n_station);
        E(3, 11) E(5, 11) E(3, 12) E(7, 11)
#undef E
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_fengine_staged(int form, int n_freq, const int8_t *in, size_t station_pitch, const int32_t *valid,
                                 size_t valid_stride, int frame0, int frames, int n_station, const float2 *tw,
                                 const float2 *wl, const float *gain, void *out, size_t out_station_stride, int tile,
                                 float2 *zs, hipStream_t) {
    if (frames <= 0 || n_station <= 0) return hipSuccess;
    const fft::StagedPlan pl = fft::staged_plan(n_freq);
    if (pl.m == 0 || form < 0 || form > 2 || (tile != 16 && tile != 32 && tile != n_freq) || frames > 65535 ||
        n_station > 65535)
        return hipErrorInvalidValue;
    if (fft::staged_one_kernel(form, pl)) {
        const int fcap = (int)(out_station_stride / (size_t)n_freq);
        // the device's one-dimensional grid, mapped by fft::staged_unit
        const unsigned X = (unsigned)(pl.m / 2 + 1), U = (unsigned)frames * (unsigned)n_station;
        auto run = [&](auto log2p_tag) {
            constexpr int L = decltype(log2p_tag)::value;
            run_grid((int)(X * U), 1, (1 << L) / 16, fft::Sizes<(1 << L)>::kLds * sizeof(float2),
                     [=](int j, int id, int, char *lds_bytes, HostSync &sync) {
                         unsigned qu, u;
                         fft::staged_unit((unsigned)id, X, U, qu, u);
                         const int bx = (int)qu, f = (int)(u % (unsigned)frames), st = (int)(u / (unsigned)frames);
                         const int frame = frame0 + f;
                         float2 *lds = reinterpret_cast<float2 *>(lds_bytes);
                         auto at = [&](int k) { return forms::spec_index(st, frame, k, n_freq, fcap, tile); };
                         const float *g = gain == nullptr ? nullptr : gain + (size_t)st * n_freq;
                         if (form == 0) {
                             float2 *o = static_cast<float2 *>(out);
                             fft::staged_frame<L>(j, bx, f, st, pl, in, station_pitch, frame0, valid, valid_stride, lds,
                                                  sync, tw, tw + n_freq, wl, [&](int k, float2 x) { o[at(k)] = x; },
                                                  [&](int k) { o[at(k)] = make_float2(0.0f, 0.0f); });
                         } else if (form == 1) {
                             uint32_t *o = static_cast<uint32_t *>(out);
                             fft::staged_frame<L>(j, bx, f, st, pl, in, station_pitch, frame0, valid, valid_stride, lds,
                                                  sync, tw, tw + n_freq, wl,
                                                  [&](int k, float2 x) { o[at(k)] = forms::pack_fp16(x, g[k]); },
                                                  [&](int k) { o[at(k)] = 0u; });
                         } else {
                             uint16_t *o = static_cast<uint16_t *>(out);
                             fft::staged_frame<L>(j, bx, f, st, pl, in, station_pitch, frame0, valid, valid_stride, lds,
                                                  sync, tw, tw + n_freq, wl,
                                                  [&](int k, float2 x) { o[at(k)] = forms::pack_int8(x, g[k]); },
                                                  [&](int k) { o[at(k)] = (uint16_t)0; });
                         }
                     });
        };
        run(std::integral_constant<int, 14>{});
        return hipSuccess;
    }
    if ((size_t)frames * (size_t)n_station > 65535u) return hipErrorInvalidValue;  // the two-kernel grids
    // stage 1, one workgroup of P/16 lanes per (k2, frame, station)
    auto sub = [&](auto log2p_tag) {
        constexpr int L = decltype(log2p_tag)::value;
        run_grid(pl.m, frames * n_station, (1 << L) / 16, fft::Sizes<(1 << L)>::kLds * sizeof(float2),
                 [=](int j, int bx, int by, char *lds, HostSync &sync) {
                     fft::staged_first_sub<L>(j, bx, by % frames, by / frames, frames, pl, in, station_pitch, frame0,
                                              valid, valid_stride, reinterpret_cast<float2 *>(lds), sync, tw,
                                              tw + n_freq, zs);
                 });
    };
    switch (pl.log2p) {
    case 11: sub(std::integral_constant<int, 11>{}); break;
    case 12: sub(std::integral_constant<int, 12>{}); break;
    case 13: sub(std::integral_constant<int, 13>{}); break;
    case 14: sub(std::integral_constant<int, 14>{}); break;
    default: return hipErrorInvalidValue;
    }
    // stage 2, every lane of the device grid
    const int lanes2 = (n_freq / 2 + 1 + 255) / 256 * 256;
    const int fcap = (int)(out_station_stride / (size_t)n_freq);
    for (int st = 0; st < n_station; ++st)
        for (int f = 0; f < frames; ++f) {
            const int frame = frame0 + f;
            for (int k = 0; k < lanes2; ++k) {
                if (form == 0) {
                    float2 *o = static_cast<float2 *>(out);
                    auto at = [&](int b) { return forms::spec_index(st, frame, b, n_freq, fcap, tile); };
                    fft::staged_r2c(k, f, st, frames, pl, frame0, valid, valid_stride, zs, wl,
                                    [&](int b, float2 x) { o[at(b)] = x; },
                                    [&](int b) { o[at(b)] = make_float2(0.0f, 0.0f); });
                } else {
                    const float *g = gain + (size_t)st * n_freq;
                    auto at = [&](int b) { return forms::spec_index(st, frame, b, n_freq, fcap, tile); };
                    if (form == 1) {
                        uint32_t *o = static_cast<uint32_t *>(out);
                        fft::staged_r2c(k, f, st, frames, pl, frame0, valid, valid_stride, zs, wl,
                                        [&](int b, float2 x) { o[at(b)] = forms::pack_fp16(x, g[b]); },
                                        [&](int b) { o[at(b)] = 0u; });
                    } else {
                        uint16_t *o = static_cast<uint16_t *>(out);
                        fft::staged_r2c(k, f, st, frames, pl, frame0, valid, valid_stride, zs, wl,
                                        [&](int b, float2 x) { o[at(b)] = forms::pack_int8(x, g[b]); },
                                        [&](int b) { o[at(b)] = (uint16_t)0; });
                    }
                }
            }
        }
    return hipSuccess;
}

hipError_t launch_unpack4(const int8_t *in, size_t in_pitch, int8_t *out, size_t out_pitch, int frames,
                          int n_freq, int n_station, hipStream_t) {
    if (frames <= 0 || n_station <= 0) return hipSuccess;
    if (n_freq % 4 != 0 || in_pitch % 4 != 0 || out_pitch % 8 != 0 ||
        in_pitch < (size_t)frames * (size_t)n_freq || out_pitch < (size_t)frames * 2u * (size_t)n_freq)
        return hipErrorInvalidValue;
    const size_t words = (size_t)frames * (size_t)n_freq / 4u;
    const size_t lanes = (words + 255) / 256 * 256;  // the device grid
    for (int st = 0; st < n_station; ++st)
        for (size_t i = 0; i < lanes; ++i)
            fft::unpack4(i, st, reinterpret_cast<const uint32_t *>(in), in_pitch / 4u,
                         reinterpret_cast<uint64_t *>(out), out_pitch / 8u, words);
    return hipSuccess;
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

hipError_t launch_fengine_form(int form, int n_freq, const int8_t *in, size_t station_pitch,
                               const int32_t *valid, size_t valid_stride, const float2 *tw,
                               const float2 *wl, const float *gain, void *out,
                               size_t out_station_stride, int tile, int frames, int n_station, hipStream_t) {
    if (frames <= 0 || n_station <= 0) return hipSuccess;
    if (form != 1 && form != 2) return hipErrorInvalidValue;
    if (tile != 16 && tile != 32 && tile != n_freq) return hipErrorInvalidValue;
    int log2_nf = 0;
    while ((1 << log2_nf) < n_freq) ++log2_nf;
    const int sh = tile == 16 ? 4 : tile == 32 ? 5 : log2_nf;
    switch (n_freq) {
#define E(M, L)                                                                                                 \
    case M << L:                                                                                                \
        return form == 1 ? fengine_mixed_form_t<M, L, 1>(in, station_pitch, valid, valid_stride, tw, wl, gain, out, \
                                                         out_station_stride, tile, frames, n_station)            \
                         : fengine_mixed_form_t<M, L, 2>(in, station_pitch, valid, valid_stride, tw, wl, gain, out, \
                                                         out_station_stride, tile, frames, n_station);
        E(3, 11) E(5, 11) E(3, 12) E(7, 11)
#undef E
#define E(L)                                                                                            \
    case 1 << L:                                                                                        \
        return form == 1 ? fengine_form_t<L, 1>(in, station_pitch, valid, valid_stride, tw, wl, gain, out, \
                                                out_station_stride, sh, frames, n_station)              \
                         : fengine_form_t<L, 2>(in, station_pitch, valid, valid_stride, tw, wl, gain, out, \
                                                out_station_stride, sh, frames, n_station);
        E(9) E(10) E(11) E(12) E(13) E(14)
#undef E
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_power(int nf, int frames, int n_station, const float2 *spec, size_t spec_station_stride,
                        const int32_t *valid, size_t valid_stride, float *power, int32_t *count, hipStream_t) {
    if (frames <= 0 || n_station <= 0 || nf <= 0) return hipSuccess;
    const int lanes = (nf + 255) / 256 * 256;  // the device grid
    for (int st = 0; st < n_station; ++st)
        for (int c = 0; c < lanes; ++c)
            fft::power_sum(c, st, nf, frames, spec, spec_station_stride, valid, valid_stride, power, count);
    return hipSuccess;
}

hipError_t launch_gains(int form, int nf, int n_station, const float *power, const int32_t *count, float *gain,
                        hipStream_t) {
    for (int st = 0; st < n_station; ++st)
        for (int c = 0; c < nf; ++c)
            gain[(size_t)st * nf + c] = fft::gain_from_power(form, power[(size_t)st * nf + c], count[st]);
    return hipSuccess;
}

hipError_t launch_fill(void *p, size_t n_words, uint32_t v, hipStream_t) {
    uint32_t *w = static_cast<uint32_t *>(p);
    for (size_t i = 0; i < n_words; ++i) w[i] = v;
    return hipSuccess;
}

hipError_t launch_noise(void *p, size_t n_words, hipStream_t) {
    uint32_t *w = static_cast<uint32_t *>(p);
    for (size_t i = 0; i < n_words; ++i) w[i] = fft::tuning_noise_word(i);
    return hipSuccess;
}

hipError_t xengine_mfma_prepare(int form, int n_station, int ch, int kb) {
    return (form >= 0 && form <= 2) && xmfma::tile_fits(n_station, ch, kb) ? hipSuccess : hipErrorInvalidValue;
}

hipError_t launch_xengine_mfma(int form, int n_station, int ch, int kb, const void *spectra, size_t ys, int nf,
                               int spec_tile, int frames, void *integ, hipStream_t) {
    if (frames <= 0) return hipSuccess;
    if (!xmfma::tile_fits(n_station, ch, kb) || nf % ch != 0) return hipErrorInvalidValue;
    if (spec_tile != ch && spec_tile != nf) return hipErrorInvalidValue;
    if (form == 0) return xmfma_f<0>(n_station, ch, kb, spectra, ys, nf, spec_tile, frames, integ);
    if (form == 1) return xmfma_f<1>(n_station, ch, kb, spectra, ys, nf, spec_tile, frames, integ);
    if (form == 2) return xmfma_f<2>(n_station, ch, kb, spectra, ys, nf, spec_tile, frames, integ);
    return hipErrorInvalidValue;
}

hipError_t launch_dump(int form, const void *integ, int n_baseline, int nf, const int2 *pairs, const float *gain,
                       uint64_t frames, float2 *out, hipStream_t) {
    const float inv = 1.0f / (float)frames;
    for (int b = 0; b < n_baseline; ++b)
        for (int c = 0; c < nf; ++c) {
            const size_t at = (size_t)b * nf + c;
            if (form == 0) {
                const float2 v = static_cast<const float2 *>(integ)[at];
                out[at] = make_float2(v.x * inv, v.y * inv);
                continue;
            }
            const double den = (double)gain[(size_t)pairs[b].x * nf + c] * (double)gain[(size_t)pairs[b].y * nf + c] *
                               (double)frames;
            double re, im;
            if (form == 1) {
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
