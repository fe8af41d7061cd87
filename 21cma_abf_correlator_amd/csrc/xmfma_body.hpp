// xmfma_body.hpp -- the MFMA X-engine of all three forms (FP32 on
// v_mfma_f32_16x16x4_f32, FP16 on v_mfma_f32_16x16x32_f16, int8 on
// v_mfma_i32_16x16x64_i8), written once for the gfx950 kernel and for a CPU
// build of the same code (one host thread per lane; the matrix instruction and
// the lane exchange are provided by an `Ops` object, so the CPU build can
// emulate them with the same lane layout).
//
// Per channel c the chunk's spectra form a real matrix G (rows x frames):
//   row 2s   = Re Y_s[t][c],  row 2s+1 = Im Y_s[t][c]   (s = input)
// and the X-engine accumulates the Gram matrix M = G G^T on 16x16 MFMA tiles,
// lower triangle of tiles only. Every visibility (j <= k) is then a 2x2 block
// of M (the reference's Re = xr*yr + xi*yi, Im = xi*yr - xr*yi with x = Y_j,
// y = Y_k; corr-cuda/src/kernels.cu:243-246):
//   Re V_jk = M[2k][2j] + M[2k+1][2j+1],   Im V_jk = M[2k][2j+1] - M[2k+1][2j]
// Products of two FP16 values are exact in FP32 and those of two int8 values
// exact in int32; FP16 words are saturated to +-65504 as they are staged
// (forms::saturate_fp16x2), so a bin that overflowed the half range counts as
// the largest finite half, never as inf; the chunk's sum is added into the integrator once per call,
// with Neumaier-compensated FP32 (FP16 form) or exactly in int32 (int8 form).
//
// The FP32 form (the drop-in's numerics, reference kernels.cu:194-265) takes the
// same path with complex64 spectra: FP32 products and FP32 sums on the matrix
// cores, a rounding order of their own within FP32 accuracy.
//
// Workgroup: CH = 16 or 32 channels, CH/2 waves, 2 channels per wave. One step
// stages 16 (FP32), 32 (FP16) or 64 (int8) frames of all inputs for the CH channels in LDS as
//   lds[channel][row][frame]  (row pitch 80 B = 64 data + 16 pad,
//                              channel pitch rows_padded*80 + 16 B)
// so that a lane's MFMA operand (16 B: 4 floats / 8 halves / 16 int8 of one
// row; the 4 floats feed four 16x16x4 products) is one 16-byte LDS read. The
// staging reads 32 B (4 channels, FP32), 16 B (4 channels, FP16) or 8 B (4
// channels, int8) per (input, frame) from the [input][frame][channel] spectra
// and transposes them in registers.
//
// MFMA lane layout assumed for the 16x16xK instructions (operands of K/4
// elements per lane):
//   A (16 x K): lane l holds row l%16, elements (l/16)*K/4 .. +K/4-1 of K
//   B (K x 16): lane l holds column l%16, the same K slice
//   D (16x16):  lane l, register v holds D[4*(l/16) + v][l%16]
// A and B take the same fragment (both are rows of G), so the K order inside a
// lane does not matter; the row/column assignment and D's layout do.
#pragma once

#include <cstddef>
#include <cstdint>

#include "forms.hpp"

namespace corr {
namespace xmfma {

/// Time-K: a staged step holds KB = 1, 2 or 4 K-extents of one MFMA (16 FP32,
/// 32 FP16 or 64 int8 frames each). One channel's staged row is then 64 * KB
/// bytes of data plus 16 of padding.
CORR_HD constexpr int row_bytes(int kb) { return 64 * kb + 16; }
CORR_HD constexpr bool kblocks_valid(int kb) { return kb == 1 || kb == 2 || kb == 4; }
constexpr int kMaxRowTiles = 6;     ///< 2*n_station <= 96
constexpr int kChanPerWave = 2;
constexpr int kMaxThreads = 1024;

/// Workgroup shape for CH channels per workgroup.
template <int CH>
struct Tile {
    static_assert(CH == 16 || CH == 32, "16 or 32 channels per workgroup");
    static constexpr int kChannels = CH;
    static constexpr int kWaves = CH / kChanPerWave;
    static constexpr int kThreads = 64 * kWaves;
    static constexpr int kQuads = CH / 4;       ///< channel quads per row
    static constexpr int kOutPitch = CH + 1;    ///< results per baseline in LDS (one spare: banks)
};

/// One lane's MFMA operand: 16 bytes.
struct Frag {
    uint32_t w[4];
};

template <class T>
struct Acc {
    T v[4];
};

/// Form-dependent constants: FP16 (spectrum word = 2 halves) and int8
/// (spectrum word = 2 int8 in 16 bits).
template <int FORM>
struct Traits;

template <>
struct Traits<0> {  // FP32: complex64 spectra, 4 x v_mfma_f32_16x16x4_f32 per K block
    using acc_t = float;
    static constexpr int kFramesPerBlock = 16;
    static constexpr int kElemBytes = 8;
};

template <>
struct Traits<1> {  // FP16
    using acc_t = float;
    static constexpr int kFramesPerBlock = 32;
    static constexpr int kElemBytes = 4;
};

template <>
struct Traits<2> {  // int8
    using acc_t = int32_t;
    static constexpr int kFramesPerBlock = 64;
    static constexpr int kElemBytes = 2;
};

CORR_HD inline int rows_padded(int n_station) { return ((2 * n_station + 15) / 16) * 16; }
CORR_HD inline int row_tiles(int n_station) { return rows_padded(n_station) / 16; }
CORR_HD inline size_t chan_pitch(int n_station, int kb = 1) {
    return (size_t)rows_padded(n_station) * (size_t)row_bytes(kb) + 16;
}

/// Dynamic LDS bytes: the staging tile of ch channels, reused for the results
/// of one call, [baseline][ch + 1] complex (the pitch one more than the
/// channels, so that one channel's baselines spread over banks).
CORR_HD inline size_t lds_bytes(int n_station, int ch, int kb = 1) {
    const size_t stage = (size_t)ch * chan_pitch(n_station, kb);
    const size_t nb = (size_t)n_station * (size_t)(n_station + 1) / 2;
    const size_t out = nb * (size_t)(ch + 1) * 8;
    return stage > out ? stage : out;
}

/// Largest LDS of one workgroup (gfx950).
constexpr size_t kLdsLimit = 160 * 1024;

/// A tile of ch channels and kb K blocks per step is available for n_station
/// inputs: 32 channels hold at most three row tiles (24 inputs), the staging
/// fits the LDS, and deeper steps (kb 2 or 4) are kept to n_station * kb <= 64,
/// where a lane stages at most two items (n_station * ch * kb items over
/// 32 * ch lanes) and so always has the next step's loads in flight.
CORR_HD inline bool tile_fits(int n_station, int ch, int kb = 1) {
    return kblocks_valid(kb) && (kb == 1 || n_station * kb <= 64) &&
           (ch == 16 || (ch == 32 && row_tiles(n_station) <= 3)) && lds_bytes(n_station, ch, kb) <= kLdsLimit;
}

/// The (row tiles, channels, K blocks) shapes the library builds: 32 channels
/// up to three row tiles; deeper steps where their staging fits the LDS at the
/// largest input count of the row tiles and a lane stages at most two items
/// (NRT * KB <= 8). Every shape tile_fits admits is among them.
template <int NRT, int CH, int KB>
constexpr bool shape_built() {
    if (CH == 32 && NRT > 3) return false;
    if (KB == 1) return true;
    return NRT * KB <= 8 && (size_t)CH * ((size_t)16 * NRT * (size_t)row_bytes(KB) + 16) <= kLdsLimit;
}

/// Workgroup -> channel tile. Consecutive workgroups are spread over the eight
/// XCDs; this sends neighbouring channel tiles (which share cache lines of the
/// spectra) to the same XCD, in dispatch order.
CORR_HD inline int channel_tile(int bx, int n_tiles) {
    if (n_tiles % 8 != 0) return bx;
    return (bx % 8) * (n_tiles / 8) + bx / 8;
}

// Neumaier's compensated sum: (s, e) += x.
CORR_HD inline void neumaier(float &s, float &e, float x) {
    const float t = s + x;
    const float as = s < 0.0f ? -s : s, ax = x < 0.0f ? -x : x;
    if (as >= ax)
        e += (s - t) + x;
    else
        e += (x - t) + s;
    s = t;
}

// The 16-byte slot of frame group o of channel ch in its staged rows is
// o ^ slot_swizzle(ch). A wave's staging stores write the same row of eight
// channels 4q + m (q < 8) at once, one channel pitch apart, and 4 pitches are
// 64 bytes modulo the 256 bytes of the LDS banks, so without the swizzle
// channels q and q + 4 fall in the same banks; the MFMA reads only need the
// same slot order for both operands of a channel, which the swizzle keeps.
CORR_HD inline int slot_swizzle(int ch) { return (ch & 16) ? 2 : 0; }

// Staging of one item: 16 B (FP16) per frame for 8 frames, or 8 B (int8) per
// frame for 16 frames, of one (input, channel quad), held in 32 registers
// between its load and its store; the store transposes them into 8 rows of
// 16 B at lds[(4q+m)][2s+p][16*o .. 16*o+15] for channel m of the quad and part
// p (0 = real, 1 = imaginary). Frames at or past `frames` load as zeros.
struct Stage {
    uint32_t w[32];
};

// Item = (input s, frame group o of 4*KB, channel quad q of CH/4), q fastest,
// so that a wave's lanes read whole rows. Frame group o is the 16-byte slot o
// of the staged rows: K block o/4, lane group o%4 of the MFMA operands.
template <int CH, int KB>
CORR_HD inline void item_coords(int item, int &q, int &o, int &s) {
    constexpr int QD = CH / 4;
    constexpr int G = 4 * KB;
    q = item % QD;
    o = (item / QD) % G;
    s = item / (G * QD);
}

template <int FORM, int CH, int KB>
struct Staging;

template <int CH, int KB>
struct Staging<0, CH, KB> {  // FP32: 4 frames x 32 B (4 complex64 channels)
    static constexpr int kRowBytes = row_bytes(KB);
    static CORR_HD inline void load(int item, const void *spectra, int nf, int fcap, int tile, int frames,
                                    int f0, int c0, Stage &st) {
        int q, o, s;
        item_coords<CH, KB>(item, q, o, s);
        // frames f0 + 4o .. +3 of channels c0 + 4q .. +3: rows `tile` elements apart
        const uint32_t *y = static_cast<const uint32_t *>(spectra) +
                            2 * forms::spec_index(s, f0 + 4 * o, c0 + 4 * q, nf, fcap, tile);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int f = f0 + 4 * o + i;
            if (f < frames) {
                const uint4 *src = reinterpret_cast<const uint4 *>(y + 2 * (size_t)i * (size_t)tile);
                const uint4 a = src[0], b = src[1];
                st.w[8 * i + 0] = a.x;  // re, im of channel 0
                st.w[8 * i + 1] = a.y;
                st.w[8 * i + 2] = a.z;  // channel 1
                st.w[8 * i + 3] = a.w;
                st.w[8 * i + 4] = b.x;  // channel 2
                st.w[8 * i + 5] = b.y;
                st.w[8 * i + 6] = b.z;  // channel 3
                st.w[8 * i + 7] = b.w;
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) st.w[8 * i + e] = 0u;
            }
        }
    }

    // Row 2s of channel m holds the real parts of frames 4o..4o+3, row 2s+1 the
    // imaginary parts: w[8i + 2m] and w[8i + 2m + 1] of frame i, no arithmetic.
    // Each word is copied into its row here, where it is consumed: with the
    // words in the rows' order from the start, the compiler copies them there
    // right after their loads, which then wait for their data at once instead
    // of staying in flight through the MFMA phase.
    static CORR_HD inline uint32_t at_store(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
        uint32_t y;
        asm volatile("v_mov_b32 %0, %1" : "=v"(y) : "v"(x));
        return y;
#else
        return x;
#endif
    }
    static CORR_HD inline void store(int item, uint8_t *lds, size_t cpitch, const Stage &st) {
        int q, o, s;
        item_coords<CH, KB>(item, q, o, s);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint4 re = make_uint4(at_store(st.w[2 * m]), at_store(st.w[8 + 2 * m]), at_store(st.w[16 + 2 * m]),
                                        at_store(st.w[24 + 2 * m]));
            const uint4 im = make_uint4(at_store(st.w[2 * m + 1]), at_store(st.w[8 + 2 * m + 1]),
                                        at_store(st.w[16 + 2 * m + 1]), at_store(st.w[24 + 2 * m + 1]));
            uint8_t *row = lds + (size_t)(4 * q + m) * cpitch + (size_t)(2 * s) * kRowBytes +
                           16 * (o ^ slot_swizzle(4 * q + m));
            *reinterpret_cast<uint4 *>(row) = re;
            *reinterpret_cast<uint4 *>(row + kRowBytes) = im;
        }
    }
};

template <int CH, int KB>
struct Staging<1, CH, KB> {  // FP16: 8 frames x 16 B (4 channels)
    static constexpr int kRowBytes = row_bytes(KB);
    static CORR_HD inline void load(int item, const void *spectra, int nf, int fcap, int tile, int frames,
                                    int f0, int c0, Stage &st) {
        int q, o, s;
        item_coords<CH, KB>(item, q, o, s);
        // frames f0 + 8o .. +7 of channels c0 + 4q .. +3: rows `tile` elements apart
        const uint32_t *y = static_cast<const uint32_t *>(spectra) +
                            forms::spec_index(s, f0 + 8 * o, c0 + 4 * q, nf, fcap, tile);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int f = f0 + 8 * o + i;
            if (f < frames) {
                const uint32_t *src = y + (size_t)i * (size_t)tile;
                const uint4 v = *reinterpret_cast<const uint4 *>(src);
                st.w[4 * i + 0] = v.x;
                st.w[4 * i + 1] = v.y;
                st.w[4 * i + 2] = v.z;
                st.w[4 * i + 3] = v.w;
            } else {
                st.w[4 * i + 0] = st.w[4 * i + 1] = st.w[4 * i + 2] = st.w[4 * i + 3] = 0u;
            }
        }
    }

    static CORR_HD inline void store(int item, uint8_t *lds, size_t cpitch, const Stage &st) {
        int q, o, s;
        item_coords<CH, KB>(item, q, o, s);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint32_t *w = st.w + m;  // w[4 i] = frame i's word of channel m
            uint4 re, im;
            re.x = (w[0] & 0xFFFFu) | (w[4] << 16);
            re.y = (w[8] & 0xFFFFu) | (w[12] << 16);
            re.z = (w[16] & 0xFFFFu) | (w[20] << 16);
            re.w = (w[24] & 0xFFFFu) | (w[28] << 16);
            im.x = (w[0] >> 16) | (w[4] & 0xFFFF0000u);
            im.y = (w[8] >> 16) | (w[12] & 0xFFFF0000u);
            im.z = (w[16] >> 16) | (w[20] & 0xFFFF0000u);
            im.w = (w[24] >> 16) | (w[28] & 0xFFFF0000u);
            // here, where the staged words are consumed, not at their load:
            // the loads of the next step stay in flight during the MFMA phase
            re.x = forms::saturate_fp16x2(re.x);
            re.y = forms::saturate_fp16x2(re.y);
            re.z = forms::saturate_fp16x2(re.z);
            re.w = forms::saturate_fp16x2(re.w);
            im.x = forms::saturate_fp16x2(im.x);
            im.y = forms::saturate_fp16x2(im.y);
            im.z = forms::saturate_fp16x2(im.z);
            im.w = forms::saturate_fp16x2(im.w);
            uint8_t *row = lds + (size_t)(4 * q + m) * cpitch + (size_t)(2 * s) * kRowBytes +
                           16 * (o ^ slot_swizzle(4 * q + m));
            *reinterpret_cast<uint4 *>(row) = re;
            *reinterpret_cast<uint4 *>(row + kRowBytes) = im;
        }
    }
};

template <int CH, int KB>
struct Staging<2, CH, KB> {  // int8: 16 frames x 8 B (4 channels)
    static constexpr int kRowBytes = row_bytes(KB);
    static CORR_HD inline void load(int item, const void *spectra, int nf, int fcap, int tile, int frames,
                                    int f0, int c0, Stage &st) {
        int q, o, s;
        item_coords<CH, KB>(item, q, o, s);
        // frames f0 + 16o .. +15 of channels c0 + 4q .. +3: rows `tile` elements apart
        const uint16_t *y = static_cast<const uint16_t *>(spectra) +
                            forms::spec_index(s, f0 + 16 * o, c0 + 4 * q, nf, fcap, tile);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int f = f0 + 16 * o + i;
            if (f < frames) {
                const uint16_t *src = y + (size_t)i * (size_t)tile;
                const uint2 v = *reinterpret_cast<const uint2 *>(src);
                st.w[2 * i + 0] = v.x;  // channels 0, 1 of the quad: re0 im0 re1 im1
                st.w[2 * i + 1] = v.y;  // channels 2, 3
            } else {
                st.w[2 * i + 0] = st.w[2 * i + 1] = 0u;
            }
        }
    }

    static CORR_HD inline void store(int item, uint8_t *lds, size_t cpitch, const Stage &st) {
        int q, o, s;
        item_coords<CH, KB>(item, q, o, s);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int sh = 8 * (2 * (m & 1) + p);  // byte of (channel m, part p)
                uint4 r;
                uint32_t *rw = &r.x;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    uint32_t acc = 0;
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        acc |= ((st.w[2 * (4 * k + b) + (m >> 1)] >> sh) & 0xFFu) << (8 * b);
                    rw[k] = acc;
                }
                uint8_t *row = lds + (size_t)(4 * q + m) * cpitch + (size_t)(2 * s + p) * kRowBytes +
                               16 * (o ^ slot_swizzle(4 * q + m));
                *reinterpret_cast<uint4 *>(row) = r;
            }
        }
    }
};

/// Whole X-engine workgroup (CH channels) for `frames` frames; NRT = row tiles,
/// KB = MFMA K blocks per staged step (time-K).
/// `ops` provides mma(Frag, Frag, Acc) and xor1(value) as wave-wide operations;
/// every lane of every wave calls them the same number of times.
template <int FORM, int NRT, int CH, int KB, class Sync, class Ops>
CORR_HD inline void xengine_mfma_block(int tid, int bx, int n_tiles, uint8_t *lds, Sync &sync, Ops &ops,
                                       const void *spectra, size_t ys, int nf, int spec_tile, int frames,
                                       int n_station, void *integ) {
    using T = Traits<FORM>;
    using acc_t = typename T::acc_t;
    using St = Staging<FORM, CH, KB>;
    constexpr int kChannels = Tile<CH>::kChannels;
    constexpr int kThreads = Tile<CH>::kThreads;
    constexpr int kOutPitch = Tile<CH>::kOutPitch;
    constexpr int kRowBytes = row_bytes(KB);
    constexpr int kFramesPerStep = T::kFramesPerBlock * KB;
    constexpr int NTP = NRT * (NRT + 1) / 2;
    const int lane = tid & 63, w = tid >> 6;
    const int c0 = channel_tile(bx, n_tiles) * kChannels;
    const int rp = NRT * 16;
    const size_t cpitch = (size_t)rp * kRowBytes + 16;
    const int n_items = n_station * kChannels * KB;  // inputs x 4*KB frame groups x CH/4 quads
    const int fcap = (int)(ys / (size_t)nf);         // frame capacity of the spectrum buffer

    // Rows 2N .. rp-1 are never staged: zero them once.
    for (int i = tid; i < kChannels * (rp - 2 * n_station) * (kRowBytes / 16); i += kThreads) {
        const int per_ch = (rp - 2 * n_station) * (kRowBytes / 16);
        const int ch = i / per_ch, rem = i - ch * per_ch;
        const int row = 2 * n_station + rem / (kRowBytes / 16), seg = rem % (kRowBytes / 16);
        *reinterpret_cast<uint4 *>(lds + (size_t)ch * cpitch + (size_t)row * kRowBytes + 16 * seg) =
            make_uint4(0u, 0u, 0u, 0u);
    }

    Acc<acc_t> acc[kChanPerWave][NTP];
#pragma unroll
    for (int ch = 0; ch < kChanPerWave; ++ch)
#pragma unroll
        for (int p = 0; p < NTP; ++p)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[ch][p].v[v] = (acc_t)0;

    // Items of this lane: tid, tid + kThreads (the second for more than 32
    // inputs, or for more than 16 / 32 inputs at KB = 2 / 4). The next step's
    // loads are issued right after this step's stores, into the same
    // registers, so they are in flight during the MFMA phase. Two staged items
    // (64 registers) fit beside the accumulators up to five row tiles (40
    // inputs, 120 accumulator registers); at six row tiles they do not, and the
    // items load each step in turn.
    constexpr int IPT = (CH * 8 * NRT * KB + kThreads - 1) / kThreads;
    constexpr bool kAhead = IPT == 1 || (IPT == 2 && NRT <= 5);
    Stage stg[IPT];
    if constexpr (kAhead) {
#pragma unroll
        for (int u = 0; u < IPT; ++u)
            if (tid + u * kThreads < n_items)
                St::load(tid + u * kThreads, spectra, nf, fcap, spec_tile, frames, 0, c0, stg[u]);
    }

    for (int f0 = 0; f0 < frames; f0 += kFramesPerStep) {
        sync();
        if constexpr (kAhead) {
#pragma unroll
            for (int u = 0; u < IPT; ++u) {
                const int item = tid + u * kThreads;
                if (item < n_items) {
                    St::store(item, lds, cpitch, stg[u]);
                    if (f0 + kFramesPerStep < frames)
                        St::load(item, spectra, nf, fcap, spec_tile, frames, f0 + kFramesPerStep, c0, stg[u]);
                }
            }
        } else {
#pragma unroll 1
            for (int item = tid; item < n_items; item += kThreads) {
                St::load(item, spectra, nf, fcap, spec_tile, frames, f0, c0, stg[0]);
                St::store(item, lds, cpitch, stg[0]);
            }
        }
        sync();
#pragma unroll
        for (int ch = 0; ch < kChanPerWave; ++ch)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
            // K block kb: slots 4 kb .. 4 kb + 3 of the channel's rows (the
            // swizzle permutes slots within a block)
            const int cl = w * kChanPerWave + ch;
            const uint8_t *base = lds + (size_t)cl * cpitch + (size_t)(lane & 15) * kRowBytes +
                                  16 * (4 * kb + ((lane >> 4) ^ slot_swizzle(cl)));
            Frag fr[NRT];
#pragma unroll
            for (int a = 0; a < NRT; ++a) {
                const uint4 v = *reinterpret_cast<const uint4 *>(base + (size_t)a * 16 * kRowBytes);
                fr[a].w[0] = v.x;
                fr[a].w[1] = v.y;
                fr[a].w[2] = v.z;
                fr[a].w[3] = v.w;
            }
            if constexpr (FORM == 0) {
                // FP32: word i of a lane's fragment is frame 4*slot + i of its
                // row, one K element of the i-th of four 16x16x4 products into
                // the same accumulator; together they cover the step's 16 frames
                int p = 0;
#pragma unroll
                for (int a = 0; a < NRT; ++a)
#pragma unroll
                    for (int b = 0; b <= a; ++b) {
                        acc[ch][p] = ops.mma_f32x4(fr[a], fr[b], acc[ch][p]);
                        ++p;
                    }
            } else {
                int p = 0;
#pragma unroll
                for (int a = 0; a < NRT; ++a)
#pragma unroll
                    for (int b = 0; b <= a; ++b) {
                        acc[ch][p] = ops.mma(fr[a], fr[b], acc[ch][p]);
                        ++p;
                    }
            }
        }
    }

    // Epilogue: 2x2 blocks of M -> visibilities, into lds[b*kOutPitch + channel].
    sync();
    acc_t *vout = reinterpret_cast<acc_t *>(lds);
    const int col = lane & 15, r0 = 4 * (lane >> 4);
#pragma unroll
    for (int ch = 0; ch < kChanPerWave; ++ch) {
        const int cl = w * kChanPerWave + ch;
        int p = 0;
#pragma unroll
        for (int a = 0; a < NRT; ++a)
#pragma unroll
            for (int b = 0; b <= a; ++b) {
#pragma unroll
                for (int v = 0; v < 4; v += 2) {
                    const acc_t m_rc = acc[ch][p].v[v];       // M[R][C]
                    const acc_t m_r1c = acc[ch][p].v[v + 1];  // M[R+1][C]
                    const acc_t n_rc = ops.xor1(m_rc);        // lane^1: M[R][C^1]
                    const acc_t n_r1c = ops.xor1(m_r1c);      // lane^1: M[R+1][C^1]
                    const int R = 16 * a + r0 + v, C = 16 * b + col;
                    const int k = R >> 1, j = C >> 1;
                    if ((col & 1) == 0 && k < n_station && j <= k) {
                        const int bl = j * n_station - (j * (j - 1)) / 2 + (k - j);
                        acc_t *o = vout + 2 * ((size_t)bl * kOutPitch + cl);
                        o[0] = m_rc + n_r1c;  // M[2k][2j] + M[2k+1][2j+1]
                        o[1] = n_rc - m_r1c;  // M[2k][2j+1] - M[2k+1][2j]
                    }
                }
                ++p;
            }
    }
    sync();

    // Flush into the integrator, channels fastest.
    const int nb = n_station * (n_station + 1) / 2;
    for (int i = tid; i < nb * kChannels; i += kThreads) {
        const int bl = i / kChannels, cl = i - bl * kChannels;
        const size_t at = (size_t)bl * (size_t)nf + (size_t)(c0 + cl);
        const size_t o = 2 * ((size_t)bl * kOutPitch + cl);
        const acc_t re = vout[o], im = vout[o + 1];
        if constexpr (FORM == 0) {
            // the call's sum added once into the complex64 row (reference
            // kernels.cu:216-264 flushes once per chunk the same way)
            float2 *g = static_cast<float2 *>(integ) + at;
            float2 cur = *g;
            cur.x += re;
            cur.y += im;
            *g = cur;
        } else if constexpr (FORM == 1) {
            float4 *g = static_cast<float4 *>(integ) + at;
            float4 cur = *g;
            neumaier(cur.x, cur.z, re);
            neumaier(cur.y, cur.w, im);
            *g = cur;
        } else {
            int2 *g = static_cast<int2 *>(integ) + at;
            int2 cur = *g;
            cur.x += re;
            cur.y += im;
            *g = cur;
        }
    }
}

}  // namespace xmfma
}  // namespace corr
