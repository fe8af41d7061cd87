// xengine_body.hpp -- the FP32 X-engine's per-workgroup algorithm, written once
// for the gfx950 kernel and for a CPU build of the same code (see fft_body.hpp).
//
// acc[b][c] += sum_t Y_j[t][c] conj(Y_k[t][c]) for baseline b = (j,k), j <= k,
// with Re = xr*yr + xi*yi and Im = xi*yr - xr*yi (reference kernels.cu:243-246),
// summed in frame order over the call's frames and added to acc once per call
// (reference kernels.cu:216-264). Each product is fused into the running sum
// (two FMAs per component), a rounding order of its own within FP32 accuracy.
//
// Workgroup: 64 channels (one per lane) x W waves. Stations are taken in blocks
// of 4; a wave owns NPAIR (jb <= kb) block pairs and a lane keeps 16 complex
// accumulators per pair, so per frame it reads 8 complex values from LDS and
// does 64 FMAs. Spectra of `frames_lds` frames x all (padded) stations x 64
// channels are staged in LDS per round: lds[(i*frames_lds + t)*64 + lane].
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace corr {
namespace xeng {

/// Launch plan of the FP32 X-engine for n_station inputs.
struct Plan {
    std::vector<int2> bpairs;  ///< (jb, kb), jb <= kb, row major
    int npair = 1;             ///< block pairs per wave
    int waves = 1;             ///< waves per workgroup
    int groups = 1;            ///< workgroups along the pair axis
    int frames_lds = 1;        ///< frames staged in LDS per round
    size_t lds_bytes = 0;
};

/// Dynamic LDS bytes of one workgroup: [npad][frames_lds][64] complex64.
CORR_HD inline size_t lds_bytes(int n_station, int frames_lds) {
    const int npad = (n_station + 3) & ~3;
    return (size_t)npad * (size_t)frames_lds * 64 * 8;
}

inline Plan make_plan(int n_station) {
    Plan p;
    const int nb = (n_station + 3) / 4;
    for (int jb = 0; jb < nb; ++jb)
        for (int kb = jb; kb < nb; ++kb) p.bpairs.push_back(make_int2(jb, kb));
    const int n = (int)p.bpairs.size();
    p.npair = n > 16 ? 2 : 1;
    const int slots = (n + p.npair - 1) / p.npair;
    p.waves = slots < 16 ? slots : 16;
    p.groups = (slots + p.waves - 1) / p.waves;
    const int npad = nb * 4;
    p.frames_lds = 49152 / (npad * 64 * 8);
    if (p.frames_lds < 1) p.frames_lds = 1;
    if (p.frames_lds > 8) p.frames_lds = 8;
    p.lds_bytes = lds_bytes(n_station, p.frames_lds);
    return p;
}

template <int NPAIR, class Sync>
CORR_HD inline void xengine_fp32_block(int tid, int nthreads, int bx, int by, float2 *sy, Sync &sync,
                                       const float2 *y, size_t ys, int nf, int frames, int n_station,
                                       const int2 *bpairs, int n_bpairs, int frames_lds,
                                       float2 *acc) {
    const int lane = tid & 63;
    const int w = tid >> 6;
    const int W = nthreads >> 6;
    const int c = bx * 64 + lane;
    const int npad = (n_station + 3) & ~3;

    int2 bp[NPAIR];
    bool act[NPAIR];
#pragma unroll
    for (int q = 0; q < NPAIR; ++q) {
        const int p = (by * W + w) * NPAIR + q;
        act[q] = p < n_bpairs;
        bp[q] = act[q] ? bpairs[p] : make_int2(0, 0);
    }

    float2 a[NPAIR][16];
#pragma unroll
    for (int q = 0; q < NPAIR; ++q)
#pragma unroll
        for (int e = 0; e < 16; ++e) a[q][e] = make_float2(0.0f, 0.0f);

    for (int f0 = 0; f0 < frames; f0 += frames_lds) {
        const int ft = frames - f0 < frames_lds ? frames - f0 : frames_lds;
        sync();
        for (int row = w; row < npad * ft; row += W) {
            const int i = row / ft, t = row - (row / ft) * ft;
            float2 val = make_float2(0.0f, 0.0f);
            if (i < n_station && c < nf) val = y[(size_t)i * ys + (size_t)(f0 + t) * nf + c];
            sy[(i * frames_lds + t) * 64 + lane] = val;
        }
        sync();
        for (int t = 0; t < ft; ++t) {
#pragma unroll
            for (int q = 0; q < NPAIR; ++q) {
                if (!act[q]) continue;
                float2 xj[4], xk[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    xj[e] = sy[((bp[q].x * 4 + e) * frames_lds + t) * 64 + lane];
                    xk[e] = sy[((bp[q].y * 4 + e) * frames_lds + t) * 64 + lane];
                }
#pragma unroll
                for (int e = 0; e < 4; ++e)
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        // each product fused into the running sum: two FMAs per
                        // component instead of a multiply, an FMA and an add
                        float2 &s = a[q][e * 4 + g];
                        s.x = __builtin_fmaf(xj[e].x, xk[g].x, __builtin_fmaf(xj[e].y, xk[g].y, s.x));
                        s.y = __builtin_fmaf(xj[e].y, xk[g].x, __builtin_fmaf(-xj[e].x, xk[g].y, s.y));
                    }
            }
        }
    }

    if (c >= nf) return;
#pragma unroll
    for (int q = 0; q < NPAIR; ++q) {
        if (!act[q]) continue;
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int jj = bp[q].x * 4 + e, kk = bp[q].y * 4 + g;
                if (jj <= kk && kk < n_station) {
                    const int b = jj * n_station - (jj * (jj - 1)) / 2 + (kk - jj);
                    float2 *o = acc + (size_t)b * nf + c;
                    float2 cur = *o;
                    cur.x += a[q][e * 4 + g].x;
                    cur.y += a[q][e * 4 + g].y;
                    *o = cur;
                }
            }
    }
}

}  // namespace xeng
}  // namespace corr
