// CPU emulation of the F-engine kernel body (csrc/fft_body.hpp): one host
// thread per lane, std::barrier for the workgroup barrier, AddressSanitizer on
// every buffer. Checks each transform against the oracle's FP64 FFT.
#include <immintrin.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

#include "emu_types.hpp"
#include "fft_body.hpp"

extern "C" void orc_fft_r2c_f64(const double *x, int L, double *out);

using namespace corr::fft;

template <int LOG2N>
static void run_block(int f, int st, const int8_t *in, size_t pitch, const int32_t *valid,
                      size_t vstride, const float2 *tw, const float2 *wl, float2 *out,
                      size_t ostride) {
    constexpr int N = 1 << LOG2N;
    constexpr int NT = N / 16;
    float2 *lds = new float2[fengine_lds_elems<LOG2N>()];  // heap, so ASan sees its bounds
    std::barrier<> bar(NT);
    std::vector<std::thread> th;
    th.reserve(NT);
    for (int j = 0; j < NT; ++j)
        th.emplace_back([=, &bar] {
            HostSync sync{&bar};
            fengine_fp32_block<LOG2N>(j, f, st, lds, sync, in, pitch, valid, vstride, tw, wl, out,
                                      ostride);
        });
    for (auto &t : th) t.join();
    delete[] lds;
}

template <int LOG2N, int FORM>
static void run_form_block(int f, int st, const int8_t *in, size_t pitch, const int32_t *valid,
                           size_t vstride, const float2 *tw, const float2 *wl, const float *gain,
                           void *out, size_t ostride, int tile_shift) {
    constexpr int N = 1 << LOG2N;
    constexpr int NT = N / 16;
    float2 *lds = new float2[fengine_lds_elems<LOG2N>()];
    std::barrier<> bar(NT);
    std::vector<std::thread> th;
    th.reserve(NT);
    for (int j = 0; j < NT; ++j)
        th.emplace_back([=, &bar] {
            HostSync sync{&bar};
            fengine_form_block<LOG2N, FORM>(j, f, st, lds, sync, in, pitch, valid, vstride, tw, wl, gain,
                                            out, ostride, tile_shift);
        });
    for (auto &t : th) t.join();
    delete[] lds;
}

// The FP16 and int8 epilogues write exactly pack(fp32 bin, gain) at the same
// index as the FP32 epilogue (the FP32 result `ref` of the same frame).
template <int LOG2N>
static int check_forms(const int8_t *in, size_t pitch, const int32_t *valid, size_t vstride,
                       const float2 *tw, const float2 *wl, const float2 *ref, size_t rstride, int nst,
                       int fr) {
    constexpr int N = 1 << LOG2N;
    std::mt19937 rng(77 + LOG2N);
    std::uniform_int_distribution<int> e(-12, 2);
    std::uniform_real_distribution<float> u(0.01f, 0.2f);
    float *g16 = new float[(size_t)nst * N], *g8 = new float[(size_t)nst * N];
    for (size_t i = 0; i < (size_t)nst * N; ++i) {
        g16[i] = std::ldexp(1.0f, e(rng));
        g8[i] = u(rng);
    }
    const size_t ostride = (size_t)fr * N;
    uint32_t *o16 = new uint32_t[ostride * nst];
    uint16_t *o8 = new uint16_t[ostride * nst];
    int bad = 0;
    for (int tile : {16, 32, N}) {  // N: the plain [station][frame][channel] order
        const int sh = tile == 16 ? 4 : tile == 32 ? 5 : LOG2N;
        for (int st = 0; st < nst; ++st)
            for (int f = 0; f < fr; ++f) {
                run_form_block<LOG2N, 1>(f, st, in, pitch, valid, vstride, tw, wl, g16, o16, ostride, sh);
                run_form_block<LOG2N, 2>(f, st, in, pitch, valid, vstride, tw, wl, g8, o8, ostride, sh);
            }
        for (int st = 0; st < nst; ++st)
            for (int f = 0; f < fr; ++f)
                for (int k = 0; k < N; ++k) {
                    const float2 y = ref[st * rstride + (size_t)f * N + k];
                    const size_t at = corr::forms::spec_index(st, f, k, N, fr, tile);  // tile-major order
                    if (o16[at] != corr::forms::pack_fp16(y, g16[(size_t)st * N + k])) ++bad;
                    if (o8[at] != corr::forms::pack_int8(y, g8[(size_t)st * N + k])) ++bad;
                }
    }
    std::printf("N=%d forms (tiles of 16, 32 and N): %d mismatches\n", N, bad);
    delete[] g16;
    delete[] g8;
    delete[] o16;
    delete[] o8;
    return bad;
}

// The host's float -> half conversion against the CPU's F16C instruction
// (round to nearest even) over normals, subnormals, overflow and specials.
__attribute__((target("f16c"))) static uint16_t f16c(float x) {
    return (uint16_t)_cvtss_sh(x, _MM_FROUND_TO_NEAREST_INT | _MM_FROUND_NO_EXC);
}

__attribute__((target("f16c"))) static float f16c_back(uint16_t h) { return _cvtsh_ss(h); }

static int check_half() {
    int bad = 0;
    std::mt19937 rng(5);
    std::vector<uint32_t> probes;
    for (uint32_t e = 90; e <= 150; ++e)  // around the subnormal, normal and overflow limits
        for (int i = 0; i < 4000; ++i) probes.push_back((e << 23) | (rng() & 0x7FFFFFu));
    for (uint32_t m = 0; m < (1u << 14); ++m) probes.push_back((112u << 23) | (m << 9));  // ties
    const uint32_t special[] = {0u, 0x7F800000u, 0x7FC00000u, 0x477FE000u, 0x477FEFFFu,
                                0x477FF000u, 0x33000000u, 0x33000001u, 0x387FC000u, 0x38800000u};
    for (uint32_t s : special) probes.push_back(s);
    for (uint32_t b : probes)
        for (uint32_t sign : {0u, 0x80000000u}) {
            float x;
            const uint32_t bits = b | sign;
            std::memcpy(&x, &bits, 4);
            const uint16_t want = f16c(x), got = corr::forms::f32_to_f16_bits(x);
            const bool nan = (want & 0x7C00u) == 0x7C00u && (want & 0x3FFu) != 0;
            if (nan ? !((got & 0x7C00u) == 0x7C00u && (got & 0x3FFu) != 0) : want != got) ++bad;
        }
    for (uint32_t h = 0; h < 0x10000u; ++h) {
        const float want = f16c_back((uint16_t)h), got = corr::forms::f16_bits_to_f32((uint16_t)h);
        if (!(want == got || (want != want && got != got))) ++bad;
        if (want == got && std::signbit(want) != std::signbit(got)) ++bad;
    }
    std::printf("half conversion: %zu probes, %d mismatches\n", probes.size() * 2 + 0x10000, bad);
    return bad;
}

template <int LOG2N>
static int check() {
    constexpr int N = 1 << LOG2N;
    const int NST = 2, FR = 3, BIT0 = 5;
    const size_t pitch = (size_t)FR * 2 * N + 64;
    std::mt19937 rng(1234 + LOG2N);
    std::uniform_int_distribution<int> d(-128, 127);
    int8_t *in = new int8_t[pitch * NST];
    for (size_t i = 0; i < pitch * NST; ++i) in[i] = (int8_t)d(rng);
    const size_t mpitch = 2;
    uint8_t *mask = new uint8_t[mpitch * NST];
    for (size_t i = 0; i < mpitch * NST; ++i) mask[i] = 0xFF;
    // station 1, frame 1 missing: bit BIT0+1
    mask[1 * mpitch + ((BIT0 + 1) >> 3)] &= (uint8_t)~(1u << ((BIT0 + 1) & 7));
    float2 *tw = new float2[N];
    float2 *wl = new float2[N];
    for (int m = 0; m < N; ++m) {
        const double a = -2.0 * M_PI * m / N, b = -M_PI * m / N;
        tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
        wl[m] = make_float2((float)std::cos(b), (float)std::sin(b));
    }
    const size_t ostride = (size_t)FR * N;
    float2 *out = new float2[ostride * NST];
    // validity expanded exactly as the validity kernel does (one lane per frame)
    int32_t *valid = new int32_t[(size_t)NST * FR];
    for (int st = 0; st < NST; ++st)
        for (int f = 0; f < FR + 5; ++f)  // lanes past `frames` must write nothing
            expand_validity(f, st, FR, mask, mpitch, BIT0, valid, FR);
    for (int st = 0; st < NST; ++st)
        for (int f = 0; f < FR; ++f)
            run_block<LOG2N>(f, st, in, pitch, valid, FR, tw, wl, out, ostride);

    int bad = 0;
    std::vector<double> x(2 * N), ref(2 * N);
    double worst = 0.0;
    for (int st = 0; st < NST; ++st)
        for (int f = 0; f < FR; ++f) {
            const float2 *y = out + st * ostride + (size_t)f * N;
            if (st == 1 && f == 1) {
                for (int k = 0; k < N; ++k)
                    if (y[k].x != 0.0f || y[k].y != 0.0f) ++bad;
                continue;
            }
            for (int n = 0; n < 2 * N; ++n) x[n] = in[st * pitch + (size_t)f * 2 * N + n];
            orc_fft_r2c_f64(x.data(), 2 * N, ref.data());
            double num = 0, den = 0;
            for (int k = 0; k < N; ++k) {
                const double dr = y[k].x - ref[2 * k], di = y[k].y - ref[2 * k + 1];
                num += dr * dr + di * di;
                den += ref[2 * k] * ref[2 * k] + ref[2 * k + 1] * ref[2 * k + 1];
            }
            const double e = std::sqrt(num / den);
            if (e > worst) worst = e;
            if (!(e < 1e-5)) ++bad;
        }
    std::printf("N=%d: worst relative RMS %.3e, %d bad\n", N, worst, bad);
    if (LOG2N == 9 || LOG2N == 14) bad += check_forms<LOG2N>(in, pitch, valid, FR, tw, wl, out, ostride, NST, FR);
    delete[] in;
    delete[] mask;
    delete[] valid;
    delete[] tw;
    delete[] wl;
    delete[] out;
    return bad;
}

// n_freq = M * 2^LOG2P: the mixed-radix body, FP32 against the oracle's FP64
// FFT (any length), the forms against pack() of the FP32 bins at every tile.
template <int M, int LOG2P, class Body>
static void run_mixed(Body body) {
    using S = MixedSizes<M, LOG2P>;
    float2 *lds = new float2[S::kLds];
    std::barrier<> bar(S::kThreads);
    std::vector<std::thread> th;
    th.reserve(S::kThreads);
    for (int j = 0; j < S::kThreads; ++j)
        th.emplace_back([=, &bar] {
            HostSync sync{&bar};
            body(j, lds, sync);
        });
    for (auto &t : th) t.join();
    delete[] lds;
}

template <int M, int LOG2P>
static int check_mixed() {
    constexpr int P = 1 << LOG2P, N = M * P;
    const int NST = 2, FR = 2, BIT0 = 3;
    const size_t pitch = (size_t)FR * 2 * N + 32;
    std::mt19937 rng(99 + N);
    std::uniform_int_distribution<int> d(-128, 127);
    int8_t *in = new int8_t[pitch * NST];
    for (size_t i = 0; i < pitch * NST; ++i) in[i] = (int8_t)d(rng);
    uint8_t mask[2] = {0xFF, (uint8_t)~(1u << (BIT0 + 1))};  // station 1, frame 1 missing
    int32_t *valid = new int32_t[(size_t)NST * FR];
    for (int st = 0; st < NST; ++st)
        for (int f = 0; f < FR; ++f) expand_validity(f, st, FR, mask, 1, BIT0, valid, FR);
    float2 *tw = new float2[N + P];
    float2 *wl = new float2[N];
    for (int m = 0; m < N; ++m) {
        const double a = -2.0 * M_PI * m / N, b = -M_PI * m / N;
        tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
        wl[m] = make_float2((float)std::cos(b), (float)std::sin(b));
    }
    for (int m = 0; m < P; ++m) {
        const double a = -2.0 * M_PI * m / P;
        tw[N + m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    const size_t ostride = (size_t)FR * N;
    float2 *out = new float2[ostride * NST];
    for (int st = 0; st < NST; ++st)
        for (int f = 0; f < FR; ++f)
            run_mixed<M, LOG2P>([=](int j, float2 *lds, HostSync &sync) {
                fengine_mixed_fp32_block<M, LOG2P>(j, f, st, lds, sync, in, pitch, valid, FR, tw, wl, out, ostride);
            });
    int bad = 0;
    double worst = 0.0;
    std::vector<double> x(2 * N), ref(2 * N);
    for (int st = 0; st < NST; ++st)
        for (int f = 0; f < FR; ++f) {
            const float2 *y = out + st * ostride + (size_t)f * N;
            if (st == 1 && f == 1) {
                for (int k = 0; k < N; ++k)
                    if (y[k].x != 0.0f || y[k].y != 0.0f) ++bad;
                continue;
            }
            for (int n = 0; n < 2 * N; ++n) x[n] = in[st * pitch + (size_t)f * 2 * N + n];
            orc_fft_r2c_f64(x.data(), 2 * N, ref.data());
            double num = 0, den = 0;
            for (int k = 0; k < N; ++k) {
                const double dr = y[k].x - ref[2 * k], di = y[k].y - ref[2 * k + 1];
                num += dr * dr + di * di;
                den += ref[2 * k] * ref[2 * k] + ref[2 * k + 1] * ref[2 * k + 1];
            }
            const double e = std::sqrt(num / den);
            if (e > worst) worst = e;
            if (!(e < 1e-5)) ++bad;
        }
    std::printf("N=%d (%d x %d): worst relative RMS %.3e, %d bad\n", N, M, P, worst, bad);
    // forms at tiles of 16, 32 and N
    float *g16 = new float[(size_t)NST * N], *g8 = new float[(size_t)NST * N];
    std::uniform_int_distribution<int> e(-12, 2);
    std::uniform_real_distribution<float> u(0.01f, 0.2f);
    for (size_t i = 0; i < (size_t)NST * N; ++i) {
        g16[i] = std::ldexp(1.0f, e(rng));
        g8[i] = u(rng);
    }
    uint32_t *o16 = new uint32_t[ostride * NST];
    uint16_t *o8 = new uint16_t[ostride * NST];
    int fbad = 0;
    for (int tile : {16, 32, N}) {
        for (int st = 0; st < NST; ++st)
            for (int f = 0; f < FR; ++f) {
                run_mixed<M, LOG2P>([=](int j, float2 *lds, HostSync &sync) {
                    fengine_mixed_form_block<M, LOG2P, 1>(j, f, st, lds, sync, in, pitch, valid, FR, tw, wl, g16,
                                                          o16, ostride, tile);
                });
                run_mixed<M, LOG2P>([=](int j, float2 *lds, HostSync &sync) {
                    fengine_mixed_form_block<M, LOG2P, 2>(j, f, st, lds, sync, in, pitch, valid, FR, tw, wl, g8,
                                                          o8, ostride, tile);
                });
            }
        for (int st = 0; st < NST; ++st)
            for (int f = 0; f < FR; ++f)
                for (int k = 0; k < N; ++k) {
                    const float2 y = out[st * ostride + (size_t)f * N + k];
                    const size_t at = corr::forms::spec_index(st, f, k, N, FR, tile);
                    if (o16[at] != corr::forms::pack_fp16(y, g16[(size_t)st * N + k])) ++fbad;
                    if (o8[at] != corr::forms::pack_int8(y, g8[(size_t)st * N + k])) ++fbad;
                }
    }
    std::printf("N=%d forms (tiles of 16, 32 and N): %d mismatches\n", N, fbad);
    delete[] g16;
    delete[] g8;
    delete[] o16;
    delete[] o8;
    delete[] in;
    delete[] valid;
    delete[] tw;
    delete[] wl;
    delete[] out;
    return bad + fbad;
}

// unpack4_word against the definition of the 4-bit staging: every byte value
// at every byte position of the word, and random words.
static int check_unpack4() {
    int bad = 0;
    auto want = [](uint32_t w) {
        uint64_t r = 0;
        for (int m = 0; m < 4; ++m) {
            const int byte = (int)((w >> (8 * m)) & 0xFFu);
            const int lo = (byte & 15) >= 8 ? (byte & 15) - 16 : (byte & 15);
            const int hi = (byte >> 4) >= 8 ? (byte >> 4) - 16 : (byte >> 4);
            r |= (uint64_t)(uint8_t)(int8_t)lo << (16 * m);
            r |= (uint64_t)(uint8_t)(int8_t)hi << (16 * m + 8);
        }
        return r;
    };
    for (int m = 0; m < 4; ++m)
        for (uint32_t v = 0; v < 256; ++v)
            if (unpack4_word(v << (8 * m)) != want(v << (8 * m))) ++bad;
    std::mt19937 rng(3);
    for (int i = 0; i < 100000; ++i) {
        const uint32_t w = rng();
        if (unpack4_word(w) != want(w)) ++bad;
    }
    std::printf("4-bit unpack: %d mismatches\n", bad);
    return bad;
}

// n_freq above 16384: the two-kernel form (its bodies in launch order) or the
// one-kernel form over the device grid, FP32 against the oracle's FP64 FFT (any
// length), one missing frame written as zeros; every bin starts as NaN, so one
// that no lane writes fails.
static int check_staged(int n, bool one_kernel) {
    const StagedPlan pl = staged_plan(n);
    const int NST = 2, FR = 2;
    const size_t pitch = (size_t)FR * 2 * n + 16;
    std::mt19937 rng(7 + n);
    std::uniform_int_distribution<int> d(-128, 127);
    std::vector<int8_t> in(pitch * NST);
    for (auto &v : in) v = (int8_t)d(rng);
    std::vector<int32_t> valid = {1, 1, 1, 0};  // station 1, frame 1 missing
    std::vector<float2> tw(n + pl.p), wl(n);
    for (int m = 0; m < n; ++m) {
        const double a = -2.0 * M_PI * m / n, b = -M_PI * m / n;
        tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
        wl[m] = make_float2((float)std::cos(b), (float)std::sin(b));
    }
    for (int m = 0; m < pl.p; ++m)
        tw[n + m] = make_float2((float)std::cos(-2.0 * M_PI * m / pl.p), (float)std::sin(-2.0 * M_PI * m / pl.p));
    std::vector<float2> zs((size_t)NST * FR * n), out((size_t)NST * FR * n);
    auto run = [&](auto tag) {
        constexpr int L = decltype(tag)::value;
        for (int st = 0; st < NST; ++st)
            for (int f = 0; f < FR; ++f)
                for (int q = 0; q <= pl.m / 2; ++q) {
                    float2 *lds = new float2[Sizes<(1 << L)>::kLds];
                    std::barrier<> bar((1 << L) / 16);
                    std::vector<std::thread> th;
                    float2 *o = out.data() + ((size_t)st * FR + f) * n;
                    for (int j = 0; j < (1 << L) / 16; ++j)
                        th.emplace_back([&, j, st, f, q, o] {
                            HostSync sync{&bar};
                            staged_frame<L>(j, q, f, st, pl, in.data(), pitch, 0, valid.data(), FR, lds, sync,
                                            tw.data(), tw.data() + n, wl.data(), [&](int b, float2 x) { o[b] = x; },
                                            [&](int b) { o[b] = make_float2(0.0f, 0.0f); });
                        });
                    for (auto &t : th) t.join();
                    delete[] lds;
                }
    };
    std::fill(out.begin(), out.end(), make_float2(std::nanf(""), 0.0f));  // every bin must be written
    auto sub = [&](auto tag) {
        constexpr int L = decltype(tag)::value;
        for (int st = 0; st < NST; ++st)
            for (int f = 0; f < FR; ++f)
                for (int k2 = 0; k2 < pl.m; ++k2) {
                    float2 *lds = new float2[Sizes<(1 << L)>::kLds];
                    std::barrier<> bar((1 << L) / 16);
                    std::vector<std::thread> th;
                    for (int j = 0; j < (1 << L) / 16; ++j)
                        th.emplace_back([&, j, st, f, k2] {
                            HostSync sync{&bar};
                            staged_first_sub<L>(j, k2, f, st, FR, pl, in.data(), pitch, 0, valid.data(), FR, lds,
                                                sync, tw.data(), tw.data() + n, zs.data());
                        });
                    for (auto &t : th) t.join();
                    delete[] lds;
                }
    };
    if (!one_kernel) {
        if (pl.log2p == 11) sub(std::integral_constant<int, 11>{});
        else if (pl.log2p == 12) sub(std::integral_constant<int, 12>{});
        else if (pl.log2p == 13) sub(std::integral_constant<int, 13>{});
        else sub(std::integral_constant<int, 14>{});
        for (int st = 0; st < NST; ++st)
            for (int f = 0; f < FR; ++f) {
                float2 *o = out.data() + ((size_t)st * FR + f) * n;
                for (int k = 0; k < (n / 2 + 1 + 255) / 256 * 256; ++k)
                    staged_r2c(k, f, st, FR, pl, 0, valid.data(), FR, zs.data(), wl.data(),
                               [&](int b, float2 x) { o[b] = x; }, [&](int b) { o[b] = make_float2(0.0f, 0.0f); });
            }
    }
    else if (pl.log2p == 11) run(std::integral_constant<int, 11>{});
    else if (pl.log2p == 12) run(std::integral_constant<int, 12>{});
    else if (pl.log2p == 13) run(std::integral_constant<int, 13>{});
    else run(std::integral_constant<int, 14>{});
    int bad = 0;
    double worst = 0.0;
    std::vector<double> x(2 * (size_t)n), ref(2 * (size_t)n);
    for (int st = 0; st < NST; ++st)
        for (int f = 0; f < FR; ++f) {
            const float2 *y = out.data() + ((size_t)st * FR + f) * n;
            if (!valid[(size_t)st * FR + f]) {
                for (int k = 0; k < n; ++k)
                    if (y[k].x != 0.0f || y[k].y != 0.0f) ++bad;
                continue;
            }
            for (int t = 0; t < 2 * n; ++t) x[t] = in[st * pitch + (size_t)f * 2 * n + t];
            orc_fft_r2c_f64(x.data(), 2 * n, ref.data());
            double num = 0, den = 0;
            for (int k = 0; k < n; ++k) {
                const double dr = y[k].x - ref[2 * k], di = y[k].y - ref[2 * k + 1];
                num += dr * dr + di * di;
                den += ref[2 * k] * ref[2 * k] + ref[2 * k + 1] * ref[2 * k + 1];
            }
            const double e = std::sqrt(num / den);
            if (e > worst) worst = e;
            if (!(e < 1e-5)) ++bad;
        }
    std::printf("N=%d staged (%d x %d, %s): worst relative RMS %.3e, %d bad\n", n, pl.m, pl.p,
                one_kernel ? "one kernel" : "two kernels", worst, bad);
    return bad;
}

// The one-kernel staged form's workgroup mapping is a bijection onto (pair,
// unit), and the pairs of one unit lie 8 ids apart outside the tail.
static int check_staged_unit() {
    int bad = 0;
    for (unsigned X = 1; X <= 17; ++X)
        for (unsigned U = 1; U <= 41; ++U) {
            std::vector<int> seen((size_t)X * U, 0);
            std::vector<long> first(U, -1);
            for (unsigned id = 0; id < X * U; ++id) {
                unsigned q, u;
                staged_unit(id, X, U, q, u);
                if (q >= X || u >= U) { ++bad; continue; }
                ++seen[(size_t)u * X + q];
                if (first[u] < 0) first[u] = id;
                else if (u < U / 8 * 8 && (id - (unsigned)first[u]) % 8 != 0) ++bad;
            }
            for (int c : seen) bad += c != 1;
        }
    std::printf("staged workgroup mapping: %d bad\n", bad);
    return bad;
}

// `fengine_emu <part> <parts>` runs every parts-th check from the part-th on, so
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
    if (mine()) bad += check_half();
    if (mine()) bad += check<9>();
    if (mine()) bad += check<10>();
    if (mine()) bad += check<11>();
    if (mine()) bad += check<12>();
    if (mine()) bad += check<13>();
    if (mine()) bad += check<14>();
    if (mine()) bad += check_mixed<3, 11>();
    if (mine()) bad += check_mixed<5, 11>();
    if (mine()) bad += check_mixed<3, 12>();
    if (mine()) bad += check_mixed<7, 11>();
    if (mine()) bad += check_unpack4();
    if (mine()) bad += check_staged(18432, false);
    if (mine()) bad += check_staged(32768, false);
    if (mine()) bad += check_staged(45056, false);
    if (mine()) bad += check_staged(67584, false);
    if (mine()) bad += check_staged(40960, false);  // m = 5
    // the one-kernel form: where the engine uses it (m = 2, 3, 4 at P = 16384),
    // and at small P and odd m, where only this check runs it
    if (mine()) bad += check_staged_unit();
    if (mine()) bad += check_staged(32768, true);
    if (mine()) bad += check_staged(49152, true);  // m = 3: a mirror pair and residue 0
    if (mine()) bad += check_staged(65536, true);  // m = 4: a mirror pair and two residues alone
    if (mine()) bad += check_staged(18432, true);
    if (mine()) bad += check_staged(67584, true);
    std::printf(bad ? "FAILED\n" : "OK\n");
    return bad ? 1 : 0;
}
