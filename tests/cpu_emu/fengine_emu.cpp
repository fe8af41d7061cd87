// CPU emulation of the F-engine kernel body (csrc/fft_body.hpp): one host
// thread per lane, std::barrier for the workgroup barrier, AddressSanitizer on
// every buffer. Checks each transform against the oracle's FP64 FFT.
#include <cmath>
#include <cstdint>
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
    float2 *lds = new float2[Sizes<N>::kLds];  // heap, so ASan sees its bounds
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
    delete[] in;
    delete[] mask;
    delete[] valid;
    delete[] tw;
    delete[] wl;
    delete[] out;
    return bad;
}

int main() {
    int bad = check<9>() + check<10>() + check<11>() + check<12>() + check<13>() + check<14>();
    std::printf(bad ? "FAILED\n" : "OK\n");
    return bad ? 1 : 0;
}
