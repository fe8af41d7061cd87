// CPU emulation of the FP32 X-engine kernel body (csrc/xengine_body.hpp) with
// the engine's own launch plan: one host thread per lane, std::barrier for the
// workgroup barrier, AddressSanitizer on every buffer. Checks against a double
// accumulation of the same products.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

#include "emu_types.hpp"
#include "xengine_body.hpp"

using namespace corr::xeng;

template <int NPAIR>
static void run(const Plan &p, int nf, int frames, int nst, const float2 *y, size_t ys, const int2 *bp,
                float2 *acc) {
    const int nthreads = 64 * p.waves;
    for (int by = 0; by < p.groups; ++by)
        for (int bx = 0; bx < (nf + 63) / 64; ++bx) {
            float2 *sy = new float2[p.lds_bytes / sizeof(float2)];
            std::barrier<> bar(nthreads);
            std::vector<std::thread> th;
            for (int t = 0; t < nthreads; ++t)
                th.emplace_back([=, &bar] {
                    HostSync sync{&bar};
                    xengine_fp32_block<NPAIR>(t, nthreads, bx, by, sy, sync, y, ys, nf, frames, nst, bp,
                                              (int)p.bpairs.size(), p.frames_lds, acc);
                });
            for (auto &t : th) t.join();
            delete[] sy;
        }
}

static int check(int nst, int nf, int frames, int cap) {
    const Plan p = make_plan(nst);
    const size_t ys = (size_t)cap * nf;
    std::mt19937 rng(nst * 131 + nf);
    std::normal_distribution<float> d(0.0f, 100.0f);
    float2 *y = new float2[ys * nst];
    for (size_t i = 0; i < ys * nst; ++i) y[i] = make_float2(d(rng), d(rng));
    const int B = nst * (nst + 1) / 2;
    float2 *acc = new float2[(size_t)B * nf];
    for (size_t i = 0; i < (size_t)B * nf; ++i) acc[i] = make_float2(1.0f, -2.0f);  // += semantics
    int2 *bp = new int2[p.bpairs.size()];
    for (size_t i = 0; i < p.bpairs.size(); ++i) bp[i] = p.bpairs[i];
    if (p.npair == 1)
        run<1>(p, nf, frames, nst, y, ys, bp, acc);
    else
        run<2>(p, nf, frames, nst, y, ys, bp, acc);
    double worst = 0;
    int b = 0;
    for (int j = 0; j < nst; ++j)
        for (int k = j; k < nst; ++k, ++b)
            for (int c = 0; c < nf; ++c) {
                double sr = 1.0, si = -2.0, mag = 0;
                for (int t = 0; t < frames; ++t) {
                    const float2 a = y[j * ys + (size_t)t * nf + c], q = y[k * ys + (size_t)t * nf + c];
                    sr += (double)a.x * q.x + (double)a.y * q.y;
                    si += (double)a.y * q.x - (double)a.x * q.y;
                    mag += std::hypot(a.x, a.y) * std::hypot(q.x, q.y);
                }
                const float2 g = acc[(size_t)b * nf + c];
                const double e = std::hypot(g.x - sr, g.y - si) / mag;
                if (e > worst) worst = e;
            }
    std::printf("N=%d nf=%d frames=%d (npair %d waves %d groups %d frames_lds %d): worst %.3e\n", nst,
                nf, frames, p.npair, p.waves, p.groups, p.frames_lds, worst);
    delete[] y;
    delete[] acc;
    delete[] bp;
    return worst < 1e-5 ? 0 : 1;
}

int main() {
    int bad = 0;
    bad += check(1, 512, 3, 4);
    bad += check(2, 512, 9, 9);
    bad += check(4, 512, 17, 20);
    bad += check(15, 512, 13, 13);
    bad += check(24, 512, 9, 16);
    bad += check(40, 512, 5, 5);
    bad += check(44, 512, 3, 7);
    bad += check(5, 130, 4, 4);  // ragged channel tile
    std::printf(bad ? "FAILED\n" : "OK\n");
    return bad ? 1 : 0;
}
