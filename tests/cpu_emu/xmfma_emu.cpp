// CPU emulation of the MFMA X-engine kernel body (csrc/xmfma_body.hpp), on
// spectra in the tile-major order of forms::spec_index, with the kernel's launch geometry: one host thread per lane, std::barrier for the
// workgroup barrier, the MFMA and lane exchange emulated with the kernel's lane
// layout (emu_mfma.hpp), AddressSanitizer on every buffer. Checks against a
// direct accumulation of the same spectra: int8 exactly (int64), FP16 against
// FP64 within FP32 rounding, integrator start values included.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

#include "emu_mfma.hpp"

using namespace corr::xmfma;

template <int FORM, int NRT, int CH, int KB>
static void run(const void *y, size_t ys, int nf, int tl, int frames, int nst, void *integ) {
    if constexpr (shape_built<NRT, CH, KB>()) {
        constexpr int T = Tile<CH>::kThreads;
        const int n_tiles = nf / CH;
        const size_t shm = lds_bytes(nst, CH, KB);
        uint8_t *lds = new uint8_t[shm];  // heap, exact size: ASan sees the bounds
        EmuMfmaShared *sh = new EmuMfmaShared;
        std::barrier<> bar(T);
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([=, &bar] {
                HostSync sync{&bar};
                EmuMfmaOps ops{t, sh, &sync};
                for (int bx = 0; bx < n_tiles; ++bx) {
                    xengine_mfma_block<FORM, NRT, CH, KB>(t, bx, n_tiles, lds, sync, ops, y, ys, nf, tl, frames,
                                                          nst, integ);
                    sync();
                }
            });
        for (auto &t : th) t.join();
        delete sh;
        delete[] lds;
    } else {
        std::printf("shape NRT=%d CH=%d KB=%d is not built\n", NRT, CH, KB);
        std::abort();
    }
}

// K blocks per staged step of the following checks (time-K).
static int g_kb = 1;

template <int FORM, int CH, int KB>
static void dispatch_kb(int nrt, int tl, const void *y, size_t ys, int nf, int frames, int nst, void *integ) {
    switch (nrt) {
    case 1: run<FORM, 1, CH, KB>(y, ys, nf, tl, frames, nst, integ); break;
    case 2: run<FORM, 2, CH, KB>(y, ys, nf, tl, frames, nst, integ); break;
    case 3: run<FORM, 3, CH, KB>(y, ys, nf, tl, frames, nst, integ); break;
    case 4: run<FORM, 4, CH, KB>(y, ys, nf, tl, frames, nst, integ); break;
    case 5: run<FORM, 5, CH, KB>(y, ys, nf, tl, frames, nst, integ); break;
    case 6: run<FORM, 6, CH, KB>(y, ys, nf, tl, frames, nst, integ); break;
    }
}

template <int FORM>
static void dispatch(int nrt, int ch, int tl, const void *y, size_t ys, int nf, int frames, int nst, void *integ) {
    if (!tile_fits(nst, ch, g_kb)) {
        std::printf("N=%d ch=%d kb=%d does not fit\n", nst, ch, g_kb);
        std::abort();
    }
#define E(C, K) \
    if (ch == C && g_kb == K) return dispatch_kb<FORM, C, K>(nrt, tl, y, ys, nf, frames, nst, integ);
    E(16, 1) E(16, 2) E(16, 4) E(32, 1) E(32, 2) E(32, 4)
#undef E
}

static int check_int8(int nst, int nf, int frames, int cap, int ch = 16, bool plain = false) {
    const int tl = plain ? nf : ch;
    const size_t ys = (size_t)cap * nf;
    std::mt19937 rng(nst * 7 + frames);
    std::uniform_int_distribution<int> d(-7, 7);
    std::vector<uint16_t> y(ys * nst);
    for (auto &v : y) v = (uint16_t)((d(rng) & 0xFF) | ((d(rng) & 0xFF) << 8));
    const int B = nst * (nst + 1) / 2;
    std::vector<int2> acc((size_t)B * nf);
    for (size_t i = 0; i < acc.size(); ++i) acc[i] = make_int2((int)(i % 97) - 40, 3 - (int)(i % 13));
    const std::vector<int2> start = acc;
    dispatch<2>(row_tiles(nst), ch, tl, y.data(), ys, nf, frames, nst, acc.data());
    int bad = 0, b = 0;
    for (int j = 0; j < nst; ++j)
        for (int k = j; k < nst; ++k, ++b)
            for (int c = 0; c < nf; ++c) {
                int64_t sr = start[(size_t)b * nf + c].x, si = start[(size_t)b * nf + c].y;
                for (int t = 0; t < frames; ++t) {
                    const uint16_t p = y[corr::forms::spec_index(j, t, c, nf, cap, tl)];
                    const uint16_t q = y[corr::forms::spec_index(k, t, c, nf, cap, tl)];
                    const int xr = (int8_t)(p & 0xFF), xi = (int8_t)(p >> 8);
                    const int yr = (int8_t)(q & 0xFF), yi = (int8_t)(q >> 8);
                    sr += xr * yr + xi * yi;
                    si += xi * yr - xr * yi;
                }
                const int2 g = acc[(size_t)b * nf + c];
                if (g.x != sr || g.y != si) ++bad;
            }
    std::printf("int8 N=%d nf=%d frames=%d ch=%d tile=%d: %d mismatches\n", nst, nf, frames, ch, tl, bad);
    std::fflush(stdout);
    return bad;
}

static int check_fp16(int nst, int nf, int frames, int cap, int ch = 16, bool plain = false) {
    const int tl = plain ? nf : ch;
    const size_t ys = (size_t)cap * nf;
    std::mt19937 rng(nst * 11 + frames);
    std::normal_distribution<float> d(0.0f, 200.0f);
    std::vector<uint32_t> y(ys * nst);
    for (auto &v : y) v = corr::forms::pack_fp16(make_float2(d(rng), d(rng)), 1.0f);
    const int B = nst * (nst + 1) / 2;
    std::vector<float4> acc((size_t)B * nf);
    for (size_t i = 0; i < acc.size(); ++i) acc[i] = float4{1000.0f, -500.0f, 0.25f, -0.125f};
    dispatch<1>(row_tiles(nst), ch, tl, y.data(), ys, nf, frames, nst, acc.data());
    double worst = 0;
    int bad = 0, b = 0;
    for (int j = 0; j < nst; ++j)
        for (int k = j; k < nst; ++k, ++b)
            for (int c = 0; c < nf; ++c) {
                double sr = 1000.25, si = -500.125, mag = 1.0;
                for (int t = 0; t < frames; ++t) {
                    const uint32_t p = y[corr::forms::spec_index(j, t, c, nf, cap, tl)];
                    const uint32_t q = y[corr::forms::spec_index(k, t, c, nf, cap, tl)];
                    const double xr = corr::forms::f16_bits_to_f32((uint16_t)p);
                    const double xi = corr::forms::f16_bits_to_f32((uint16_t)(p >> 16));
                    const double yr = corr::forms::f16_bits_to_f32((uint16_t)q);
                    const double yi = corr::forms::f16_bits_to_f32((uint16_t)(q >> 16));
                    sr += xr * yr + xi * yi;
                    si += xi * yr - xr * yi;
                    mag += std::hypot(xr, xi) * std::hypot(yr, yi);
                }
                const float4 g = acc[(size_t)b * nf + c];
                const double e = std::hypot((double)g.x + g.z - sr, (double)g.y + g.w - si) / mag;
                if (e > worst) worst = e;
                if (!(e < 2e-6)) ++bad;
            }
    std::printf("fp16 N=%d nf=%d frames=%d ch=%d tile=%d: worst %.3e, %d bad\n", nst, nf, frames, ch, tl, worst,
                bad);
    std::fflush(stdout);
    return bad;
}

// FP32 form: complex64 spectra in the plain order (the only order its F-engine
// writes), against FP64 within FP32 rounding, integrator start values included.
static int check_fp32(int nst, int nf, int frames, int cap, int ch = 16) {
    const size_t ys = (size_t)cap * nf;
    std::mt19937 rng(nst * 13 + frames);
    std::normal_distribution<float> d(0.0f, 3000.0f);
    std::vector<float2> y(ys * nst);
    for (auto &v : y) v = make_float2(d(rng), d(rng));
    const int B = nst * (nst + 1) / 2;
    std::vector<float2> acc((size_t)B * nf);
    for (size_t i = 0; i < acc.size(); ++i) acc[i] = make_float2(1000.0f, -500.0f);
    dispatch<0>(row_tiles(nst), ch, nf, y.data(), ys, nf, frames, nst, acc.data());
    double worst = 0;
    int bad = 0, b = 0;
    for (int j = 0; j < nst; ++j)
        for (int k = j; k < nst; ++k, ++b)
            for (int c = 0; c < nf; ++c) {
                double sr = 1000.0, si = -500.0, mag = 1.0;
                for (int t = 0; t < frames; ++t) {
                    const float2 p = y[(size_t)j * ys + (size_t)t * nf + c];
                    const float2 q = y[(size_t)k * ys + (size_t)t * nf + c];
                    sr += (double)p.x * q.x + (double)p.y * q.y;
                    si += (double)p.y * q.x - (double)p.x * q.y;
                    mag += std::hypot((double)p.x, (double)p.y) * std::hypot((double)q.x, (double)q.y);
                }
                const float2 g = acc[(size_t)b * nf + c];
                const double e = std::hypot((double)g.x - sr, (double)g.y - si) / mag;
                if (e > worst) worst = e;
                if (!(e < 2e-6)) ++bad;
            }
    std::printf("fp32 N=%d nf=%d frames=%d ch=%d: worst %.3e, %d bad\n", nst, nf, frames, ch, worst, bad);
    std::fflush(stdout);
    return bad;
}

// FP16 words of +-inf (bins past the half range) are read as +-65504.
static int check_fp16_saturation() {
    const int nst = 3, nf = 48, frames = 9, cap = 9;
    const size_t ys = (size_t)cap * nf;
    std::vector<uint32_t> y(ys * nst);
    std::vector<float2> v(ys * nst);
    std::mt19937 rng(5);
    std::normal_distribution<float> d(0.0f, 100.0f);
    for (size_t i = 0; i < y.size(); ++i) {
        float2 x = make_float2(d(rng), d(rng));
        if (i % 7 == 3) x.x = 2.0e5f;    // rounds to +inf in FP16
        if (i % 11 == 5) x.y = -9.0e4f;  // -inf
        y[i] = corr::forms::pack_fp16(x, 1.0f);
        const float sr = x.x > 65504.0f ? 65504.0f : x.x, si = x.y < -65504.0f ? -65504.0f : x.y;
        v[i] = make_float2(corr::forms::f16_bits_to_f32(corr::forms::f32_to_f16_bits(sr)),
                           corr::forms::f16_bits_to_f32(corr::forms::f32_to_f16_bits(si)));
    }
    const int B = nst * (nst + 1) / 2;
    std::vector<float4> acc((size_t)B * nf, float4{0.0f, 0.0f, 0.0f, 0.0f});
    dispatch<1>(row_tiles(nst), 16, 16, y.data(), ys, nf, frames, nst, acc.data());
    int bad = 0, b = 0;
    for (int j = 0; j < nst; ++j)
        for (int k = j; k < nst; ++k, ++b)
            for (int c = 0; c < nf; ++c) {
                double sr = 0, si = 0, mag = 1.0;
                for (int t = 0; t < frames; ++t) {
                    const float2 p = v[corr::forms::spec_index(j, t, c, nf, cap, 16)];
                    const float2 q = v[corr::forms::spec_index(k, t, c, nf, cap, 16)];
                    sr += (double)p.x * q.x + (double)p.y * q.y;
                    si += (double)p.y * q.x - (double)p.x * q.y;
                    mag += std::hypot((double)p.x, (double)p.y) * std::hypot((double)q.x, (double)q.y);
                }
                const float4 g = acc[(size_t)b * nf + c];
                const double e = std::hypot((double)g.x + g.z - sr, (double)g.y + g.w - si) / mag;
                if (!(e < 2e-6)) ++bad;  // also counts inf and NaN
            }
    std::printf("fp16 saturation: %d bad\n", bad);
    std::fflush(stdout);
    return bad;
}

// `xmfma_emu <part> <parts>` runs every parts-th check from the part-th on, so
// that the checks can run in parallel processes; no arguments run them all.
static int g_part = 0, g_parts = 1, g_idx = 0;
static bool mine() { return (g_idx++ % g_parts) == g_part; }

int main(int argc, char **argv) {
    if (argc == 3) {
        g_part = std::atoi(argv[1]);
        g_parts = std::atoi(argv[2]);
        if (g_parts < 1 || g_part < 0 || g_part >= g_parts) return 2;
    }
    int bad = 0;
    if (mine()) bad += check_fp32(1, 48, 3, 4);
    if (mine()) bad += check_fp32(5, 256, 37, 40);
    if (mine()) bad += check_fp32(24, 128, 20, 20);
    if (mine()) bad += check_fp32(24, 96, 33, 40, 32);
    if (mine()) bad += check_fp32(40, 48, 17, 17);  // five row tiles, two items per lane, prefetched
    if (mine()) bad += check_fp32(48, 48, 16, 16);  // six row tiles, items in turn
    if (mine()) bad += check_fp16_saturation();
    // nf 256: 16 channel tiles (XCD mapping permutes); nf 128: 8, nf 48: 3 (identity)
    if (mine()) bad += check_int8(1, 48, 3, 4);
    if (mine()) bad += check_int8(4, 256, 70, 72);
    if (mine()) bad += check_int8(9, 48, 130, 130);
    if (mine()) bad += check_int8(24, 128, 64, 64);
    if (mine()) bad += check_int8(30, 48, 20, 20);
    if (mine()) bad += check_int8(36, 48, 70, 70);  // five row tiles, two items per lane, prefetched
    if (mine()) bad += check_int8(48, 48, 65, 66);
    if (mine()) bad += check_fp16(1, 48, 1, 2);
    if (mine()) bad += check_fp16(5, 256, 33, 40);
    if (mine()) bad += check_fp16(8, 48, 64, 64);
    if (mine()) bad += check_fp16(24, 128, 20, 20);
    if (mine()) bad += check_fp16(40, 48, 70, 70);  // five row tiles, two items per lane, prefetched
    // 32 channels per workgroup (up to 24 inputs): nf 512 = 16 tiles, nf 96 = 3
    if (mine()) bad += check_int8(3, 96, 70, 70, 32);
    if (mine()) bad += check_int8(24, 96, 40, 40, 32);
    if (mine()) bad += check_fp16(13, 96, 33, 33, 32);
    if (mine()) bad += check_fp16(8, 512, 20, 20, 32);
    // the plain [input][frame][channel] order (a spectrum tile of n_freq)
    if (mine()) bad += check_int8(9, 96, 70, 72, 16, true);
    if (mine()) bad += check_int8(24, 96, 40, 40, 32, true);
    if (mine()) bad += check_fp16(5, 512, 33, 40, 16, true);
    if (mine()) bad += check_fp16(24, 96, 40, 40, 32, true);
    // deeper staged steps (time-K): two and four K extents of frames, ragged
    // frame counts, one and two items per lane, both widths and orders
    g_kb = 2;
    if (mine()) bad += check_fp32(13, 48, 70, 72);
    if (mine()) bad += check_fp32(16, 96, 33, 33, 32);
    if (mine()) bad += check_fp16(5, 256, 150, 150);
    if (mine()) bad += check_fp16(24, 48, 70, 70);   // two items per lane
    if (mine()) bad += check_fp16(16, 96, 90, 90, 32, true);
    if (mine()) bad += check_int8(9, 48, 200, 200);
    if (mine()) bad += check_int8(30, 48, 140, 140);  // four row tiles, two items per lane
    g_kb = 4;
    if (mine()) bad += check_fp32(8, 96, 70, 70, 32);
    if (mine()) bad += check_fp16(3, 48, 130, 130);
    if (mine()) bad += check_fp16(16, 48, 140, 140, 16, true);  // two row tiles, two items per lane
    if (mine()) bad += check_int8(6, 96, 300, 300, 32);
    g_kb = 1;
    std::printf(bad ? "FAILED\n" : "OK\n");
    return bad ? 1 : 0;
}
