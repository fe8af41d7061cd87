// fft_body.hpp -- the F-engine's per-workgroup algorithm, written once for the
// gfx950 kernel and for a CPU build of the same code (one host thread per lane,
// a barrier object for the workgroup barrier), so that every index of the
// kernel can be exercised with bounds checking on the CPU.
//
// The includer provides float2/char2/make_float2 (HIP's vector types on the
// device) and defines CORR_HD (__host__ __device__ under hipcc, empty on the CPU).
//
// Algorithm (the reference's cuFFT R2C, corr-cuda/src/kernels.cu:578-592, 832-835,
// restated from the mathematics): the 2*N real samples of a frame are read as N
// complex values z[m] = x[2m] + i x[2m+1]; an N-point complex FFT is done as a
// Stockham autosort FFT (radix-16 passes, one last radix-2/4/8 pass), each pass a
// register-resident radix-R DFT between LDS exchanges; then the N retained R2C bins
//   X[k] = (Z[k] + conj(Z[N-k]))/2 + W^k (Z[k] - conj(Z[N-k]))/(2i), W = exp(-i pi/N)
// are formed (the Nyquist bin k = N is never computed).
#pragma once

#include <cstddef>
#include <cstdint>
#include <type_traits>

#include "forms.hpp"

// Every function of the F-engine body is inlined into its kernel: an outlined
// call would pass the LDS array as a generic pointer (address-space casts on
// every access) and pay the call's register save and restore. The CPU build
// leaves inlining to the compiler.
#ifndef CORR_FI
#if defined(__HIP_DEVICE_COMPILE__)
#define CORR_FI __attribute__((always_inline)) inline
#else
#define CORR_FI inline
#endif
#endif

namespace corr {
namespace fft {

#if defined(__HIP_DEVICE_COMPILE__)
// A raw buffer resource over [base, base + 2 GiB). A load or store through it
// takes the lane's 32-bit byte offset in a vector register and everything else
// (base, run offset) in scalar registers, so it needs no vector address
// arithmetic; a global access with a per-lane 64-bit address costs one or two
// VALU instructions, which the F-engine (VALU-bound) pays on every access.
// 0x00020000: the gfx9 descriptor word 3 (32-bit data format, no swizzle).
CORR_FI __device__ __amdgpu_buffer_rsrc_t raw_rsrc(const void *base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, 0x7FFFFFFF, 0x00020000);
}
#endif

// Word i of the autotuner's synthetic input: four int8 samples in [-64, 63]
// from a hash of i (an integer mix in the manner of MurmurHash3's finalizer).
// A transform of noise draws about the power that real samples do, and the
// part's clock follows its power; one of zeros draws less.
CORR_HD CORR_FI uint32_t tuning_noise_word(uint64_t i) {
    uint32_t h = (uint32_t)i ^ (uint32_t)(i >> 32) * 0x85EBCA6Bu;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    uint32_t w = 0;
    for (int k = 0; k < 4; ++k) w |= (uint32_t)(uint8_t)((int8_t)(uint8_t)(h >> (8 * k)) >> 1) << (8 * k);
    return w;
}

CORR_HD CORR_FI float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
CORR_HD CORR_FI float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
CORR_HD CORR_FI float2 cmul(float2 a, float2 w) {
    return make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}

// cos(2 pi m / 32) for m = 0..8, correctly rounded to float.
CORR_HD constexpr float cos_oct(int m) {
    return m == 0   ? 1.0f
           : m == 1 ? 0.98078528040323044913f
           : m == 2 ? 0.92387953251128675613f
           : m == 3 ? 0.83146961230254523708f
           : m == 4 ? 0.70710678118654752440f
           : m == 5 ? 0.55557023301960222474f
           : m == 6 ? 0.38268343236508977173f
           : m == 7 ? 0.19509032201612826785f
                    : 0.0f;
}

CORR_HD constexpr float cos32(int m) {
    return m <= 8 ? cos_oct(m) : m <= 16 ? -cos_oct(16 - m) : m <= 24 ? -cos_oct(m - 16) : cos_oct(32 - m);
}

// x * exp(-2 pi i M / 32) for a compile-time M in [0, 16).
template <int M>
CORR_HD CORR_FI float2 rot32(float2 x) {
    if constexpr (M == 0) {
        return x;
    } else if constexpr (M == 8) {  // -i
        return make_float2(x.y, -x.x);
    } else if constexpr (M == 4) {  // (1 - i)/sqrt2
        constexpr float r = cos_oct(4);
        return make_float2((x.x + x.y) * r, (x.y - x.x) * r);
    } else if constexpr (M == 12) {  // (-1 - i)/sqrt2
        constexpr float r = cos_oct(4);
        return make_float2((x.y - x.x) * r, -(x.x + x.y) * r);
    } else {
        constexpr float c = cos32(M);
        constexpr float s = cos32((M + 24) & 31);  // sin(2 pi M / 32)
        return make_float2(x.x * c + x.y * s, x.y * c - x.x * s);
    }
}

// x * exp(-2 pi i M / 32) for a compile-time M in [0, 32).
template <int M>
CORR_HD CORR_FI float2 rot32_any(float2 x) {
    if constexpr (M >= 16) {
        const float2 y = rot32<M - 16>(x);
        return make_float2(-y.x, -y.y);
    } else {
        return rot32<M>(x);
    }
}

// One radix-2 DIF stage of an in-register DFT of size R: blocks of 2*H.
template <int R, int H, int BASE, int K>
CORR_HD CORR_FI void dif_butterflies(float2 *v) {
    if constexpr (BASE < R) {
        if constexpr (K < H) {
            const float2 a = v[BASE + K], b = v[BASE + K + H];
            v[BASE + K] = cadd(a, b);
            v[BASE + K + H] = rot32<K * (32 / (2 * H))>(csub(a, b));  // W_{2H}^K
            dif_butterflies<R, H, BASE, K + 1>(v);
        } else {
            dif_butterflies<R, H, BASE + 2 * H, 0>(v);
        }
    }
}

template <int R, int H>
CORR_HD CORR_FI void dif_stages(float2 *v) {
    if constexpr (H >= 1) {
        dif_butterflies<R, H, 0, 0>(v);
        dif_stages<R, H / 2>(v);
    }
}

template <int R>
CORR_HD constexpr int bitrev(int i) {
    int r = 0;
    for (int b = 1; b < R; b <<= 1) {
        r = (r << 1) | (i & 1);
        i >>= 1;
    }
    return r;
}

template <int R, int I>
CORR_HD CORR_FI void unscramble(float2 *v) {
    if constexpr (I < R) {
        constexpr int J = bitrev<R>(I);
        if constexpr (J > I) {
            const float2 t = v[I];
            v[I] = v[J];
            v[J] = t;
        }
        unscramble<R, I + 1>(v);
    }
}

// In-place 16-point DFT as 4 x 4 (Cooley-Tukey, n = n1 + 4 n2, k = 4 k1 + k2):
// four radix-4 DFTs over n2, the twiddles W16^(n1 k2), four radix-4 DFTs over
// n1. Every constant twiddle is applied in a scaled form that an FMA absorbs:
// W16^1 = C (1 - i T) and W16^3 = C (T - i) (C = cos(pi/8), T = tan(pi/8)) share
// the factor C, W16^2 = R (1 - i) and W16^6 = R (-1 - i) (R = cos(pi/4)) the
// factor R, so each column's products are formed unscaled (additions, or one
// FMA with T per component) and the common factor enters in the column's final
// FMAs. 144 floating-point instructions where the radix-2 form takes 162.
CORR_HD CORR_FI void dft16_fma(float2 *v) {
    constexpr float C = 0.92387953251128675613f;  // cos(pi/8)
    constexpr float T = 0.41421356237309504880f;  // tan(pi/8)
    constexpr float R = 0.70710678118654752440f;  // cos(pi/4)
    float2 y[4][4];  // y[n1][k2]: radix-4 DFT of v[n1 + 4 n2] over n2
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
        const float2 a0 = v[n1], a1 = v[n1 + 4], a2 = v[n1 + 8], a3 = v[n1 + 12];
        const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = csub(a1, a3);
        y[n1][0] = cadd(t0, t2);
        y[n1][2] = csub(t0, t2);
        y[n1][1] = make_float2(t1.x + t3.y, t1.y - t3.x);  // t1 - i t3
        y[n1][3] = make_float2(t1.x - t3.y, t1.y + t3.x);  // t1 + i t3
    }
    // Column k2: X[4 k1 + k2] = radix-4 DFT over n1 of y[n1][k2] W16^(n1 k2).
    {  // k2 = 0: no twiddles
        const float2 t0 = cadd(y[0][0], y[2][0]), t1 = csub(y[0][0], y[2][0]);
        const float2 t2 = cadd(y[1][0], y[3][0]), t3 = csub(y[1][0], y[3][0]);
        v[0] = cadd(t0, t2);
        v[8] = csub(t0, t2);
        v[4] = make_float2(t1.x + t3.y, t1.y - t3.x);
        v[12] = make_float2(t1.x - t3.y, t1.y + t3.x);
    }
    {  // k2 = 1: y1 W^1 = C (x + T y, y - T x), y2 W^2 = R (x + y, y - x), y3 W^3 = C (y + T x, T y - x)
        const float2 Y0 = y[0][1], Y1 = y[1][1], Y2 = y[2][1], Y3 = y[3][1];
        const float qx = Y2.x + Y2.y, qy = Y2.y - Y2.x;
        const float2 t0 = make_float2(__builtin_fmaf(R, qx, Y0.x), __builtin_fmaf(R, qy, Y0.y));
        const float2 t1 = make_float2(__builtin_fmaf(-R, qx, Y0.x), __builtin_fmaf(-R, qy, Y0.y));
        const float2 p1 = make_float2(__builtin_fmaf(T, Y1.y, Y1.x), __builtin_fmaf(-T, Y1.x, Y1.y));
        const float2 p3 = make_float2(__builtin_fmaf(T, Y3.x, Y3.y), __builtin_fmaf(T, Y3.y, -Y3.x));
        const float2 u = cadd(p1, p3), w = csub(p1, p3);  // t2 = C u, t3 = C w
        v[1] = make_float2(__builtin_fmaf(C, u.x, t0.x), __builtin_fmaf(C, u.y, t0.y));
        v[9] = make_float2(__builtin_fmaf(-C, u.x, t0.x), __builtin_fmaf(-C, u.y, t0.y));
        v[5] = make_float2(__builtin_fmaf(C, w.y, t1.x), __builtin_fmaf(-C, w.x, t1.y));
        v[13] = make_float2(__builtin_fmaf(-C, w.y, t1.x), __builtin_fmaf(C, w.x, t1.y));
    }
    {  // k2 = 2: y1 W^2 = R (x + y, y - x), y2 W^4 = -i y2, y3 W^6 = R (y - x, -(x + y))
        const float2 Y0 = y[0][2], Y1 = y[1][2], Y2 = y[2][2], Y3 = y[3][2];
        const float2 t0 = make_float2(Y0.x + Y2.y, Y0.y - Y2.x), t1 = make_float2(Y0.x - Y2.y, Y0.y + Y2.x);
        const float q2x = Y1.x + Y1.y, q2y = Y1.y - Y1.x;
        const float s3 = Y3.x + Y3.y, q6x = Y3.y - Y3.x;
        const float2 u = make_float2(q2x + q6x, q2y - s3), w = make_float2(q2x - q6x, q2y + s3);  // t2 = R u, t3 = R w
        v[2] = make_float2(__builtin_fmaf(R, u.x, t0.x), __builtin_fmaf(R, u.y, t0.y));
        v[10] = make_float2(__builtin_fmaf(-R, u.x, t0.x), __builtin_fmaf(-R, u.y, t0.y));
        v[6] = make_float2(__builtin_fmaf(R, w.y, t1.x), __builtin_fmaf(-R, w.x, t1.y));
        v[14] = make_float2(__builtin_fmaf(-R, w.y, t1.x), __builtin_fmaf(R, w.x, t1.y));
    }
    {  // k2 = 3: y1 W^3 = C (y + T x, T y - x), y2 W^6 = R (y - x, -(x + y)), y3 W^9 = -C (x + T y, y - T x)
        const float2 Y0 = y[0][3], Y1 = y[1][3], Y2 = y[2][3], Y3 = y[3][3];
        const float qx = Y2.y - Y2.x, qy = -(Y2.x + Y2.y);
        const float2 t0 = make_float2(__builtin_fmaf(R, qx, Y0.x), __builtin_fmaf(R, qy, Y0.y));
        const float2 t1 = make_float2(__builtin_fmaf(-R, qx, Y0.x), __builtin_fmaf(-R, qy, Y0.y));
        const float2 p3 = make_float2(__builtin_fmaf(T, Y1.x, Y1.y), __builtin_fmaf(T, Y1.y, -Y1.x));
        const float2 p1 = make_float2(__builtin_fmaf(T, Y3.y, Y3.x), __builtin_fmaf(-T, Y3.x, Y3.y));
        const float2 u = csub(p3, p1), w = cadd(p3, p1);  // t2 = C u, t3 = C w
        v[3] = make_float2(__builtin_fmaf(C, u.x, t0.x), __builtin_fmaf(C, u.y, t0.y));
        v[11] = make_float2(__builtin_fmaf(-C, u.x, t0.x), __builtin_fmaf(-C, u.y, t0.y));
        v[7] = make_float2(__builtin_fmaf(C, w.y, t1.x), __builtin_fmaf(-C, w.x, t1.y));
        v[15] = make_float2(__builtin_fmaf(-C, w.y, t1.x), __builtin_fmaf(C, w.x, t1.y));
    }
}

// In-place DFT of R (power of two, <= 32) values, natural order in and out:
// v[k] := sum_r v[r] exp(-2 pi i r k / R).
template <int R>
CORR_HD CORR_FI void dft(float2 *v) {
    if constexpr (R == 16) {
        dft16_fma(v);
    } else {
        dif_stages<R, R / 2>(v);
        unscramble<R, 0>(v);
    }
}

// LDS layouts of the exchange buffer. A pass whose lanes write with a stride
// (the first pass, stride 16, and the NS = 16 pass, runs of 16) writes the
// padded layout, one spare float2 after every 16, so that a wave's writes spread
// over all banks; a pass whose lanes write contiguous runs of 64 or more (NS >=
// 64) writes the plain layout. Every read uses the layout its data was written
// in. Offsets that are multiples of 16 elements map linearly in both layouts,
// so each access is one base address plus a compile-time offset.
CORR_HD CORR_FI int pad16(int a) { return a + (a >> 4); }

template <bool PAD>
CORR_HD CORR_FI int lds_index(int a) { return PAD ? pad16(a) : a; }

// Elements of an offset that is a multiple of 16, in the layout.
template <bool PAD>
CORR_HD constexpr int lds_span(int a16) { return PAD ? a16 + a16 / 16 : a16; }

// The layout written by a pass whose butterflies combine NS points.
CORR_HD constexpr bool pad_for(int ns) { return ns < 64; }

template <int N>
struct Sizes {
    static constexpr int kThreads = N / 16;
    static constexpr int kLds = N + N / 16;  // float2 elements
};

// Largest power of two <= r (r >= 1), at compile time.
CORR_HD constexpr int high_bit(int r) { return r >= 16 ? 16 : r >= 8 ? 8 : r >= 4 ? 4 : r >= 2 ? 2 : 1; }

// The twiddles w^r, r = 1..R-1, of one butterfly group, w = exp(-2 pi i m/N)
// with m = k*TWSTEP: the powers of two w, w^2, w^4, w^8 are read from the
// FP64-built table, every other power is the product of two earlier ones
// (w^r = w^(2^h) w^(r - 2^h), at most four rounded factors). Table reads are
// the F-engine's costliest memory operation (a wave's lanes read scattered
// entries), so this trades eleven of them per radix-16 group for complex products.
template <int R, int TWSTEP>
CORR_HD CORR_FI void group_twiddle_reads(const float2 *tw, int k, float2 (&w)[R - 1]) {
#pragma unroll
    for (int r = 1; r < R; ++r) {
        if ((r & (r - 1)) == 0)
            w[r - 1] = tw[(unsigned)(r * k * TWSTEP)];
    }
}

template <int R>
CORR_HD CORR_FI void group_twiddle_products(float2 (&w)[R - 1]) {
#pragma unroll
    for (int r = 3; r < R; ++r) {
        if ((r & (r - 1)) != 0) {
            const int h = high_bit(r);
            w[r - 1] = cmul(w[h - 1], w[r - h - 1]);
        }
    }
}

template <int R, int TWSTEP>
CORR_HD CORR_FI void group_twiddles(const float2 *tw, int k, float2 (&w)[R - 1]) {
    group_twiddle_reads<R, TWSTEP>(tw, k, w);
    group_twiddle_products<R>(w);
}

// The twiddles of the second pass (R = 16, NS = 16) depend on the lane's group
// k = j mod 16 alone: W_256^(k r) for r = 1..15, 240 values. The power-of-two
// F-engine copies them from the FP64-built table into LDS, behind its exchange
// buffer, before its first barrier; the pass then reads its 15 from one base
// address, in place of 4 scattered table reads and 11 complex products per lane.
constexpr int kTw16Elems = 16 * 15;

template <int N>
CORR_HD CORR_FI void fill_tw16(int j, float2 *tab, const float2 *tw) {
    for (int i = j; i < kTw16Elems; i += N / 16) {
        const int k = i / 15, r = i - 15 * k + 1;
        tab[i] = tw[(unsigned)(k * r * (N / 256))];
    }
}

CORR_HD CORR_FI void read_tw16(const float2 *tab, int j, float2 (&w)[15]) {
    const float2 *t = tab + 15 * (j & 15);
#pragma unroll
    for (int r = 0; r < 15; ++r) w[r] = t[r];
}

// One Stockham pass of radix R over N points, NS = points already combined,
// in place in `lds` (barrier between all reads and any write). Thread j does
// 16/R butterflies jj = j + q*N/16. The pass's twiddles are requested before
// its LDS reads and barrier, so their cache latency overlaps the exchange.
// With TAB, the second pass (NS = 16) reads its twiddles from `tab` (fill_tw16).
template <int N, int R, int NS, bool TAB, class Sync>
CORR_HD CORR_FI void stockham_pass(float2 (&v)[16], float2 *lds, const float2 *tw, const float2 *tab, int j,
                                   Sync &sync) {
    constexpr int NT = N / 16;
    constexpr int NB = 16 / R;
    constexpr int STRIDE = N / R;
    constexpr int TWSTEP = N / (NS * R);
    constexpr bool PIN = pad_for(NS / 16 > 0 ? NS / 16 : 1);  // layout of the previous pass's writes
    constexpr bool POUT = pad_for(NS);
    static_assert(STRIDE % 16 == 0 && NS % 16 == 0, "offsets are whole 16-element runs");
    float2 w[NB][R - 1];
    if constexpr (NS * R == N && NB > 1) {
        // Last pass: k = j + q*N/16 exactly, so group q's twiddles are group 0's
        // times exp(-2 pi i r q/16), constants.
        group_twiddles<R, TWSTEP>(tw, j, w[0]);
#pragma unroll
        for (int q = 1; q < NB; ++q)
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const int m = (2 * r * q) & 31;  // exp(-2 pi i m/32)
                w[q][r - 1] = cmul(w[0][r - 1], make_float2(cos32(m), -cos32((m + 24) & 31)));
            }
    } else if constexpr (TAB && NS == 16 && R == 16) {
        read_tw16(tab, j, w[0]);
    } else {
#pragma unroll
        for (int q = 0; q < NB; ++q) group_twiddles<R, TWSTEP>(tw, (j + q * NT) & (NS - 1), w[q]);
    }
    {
        const float2 *src = lds + lds_index<PIN>(j);
#pragma unroll
        for (int q = 0; q < NB; ++q)
#pragma unroll
            for (int r = 0; r < R; ++r) v[q * R + r] = src[lds_span<PIN>(q * NT + r * STRIDE)];
    }
    sync();
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int jj = j + q * NT;
        const int k = jj & (NS - 1);
#pragma unroll
        for (int r = 1; r < R; ++r) v[q * R + r] = cmul(v[q * R + r], w[q][r - 1]);
        dft<R>(&v[q * R]);
        float2 *dst = lds + lds_index<POUT>((jj - k) * R + k);
#pragma unroll
        for (int r = 0; r < R; ++r) dst[lds_span<POUT>(r * NS)] = v[q * R + r];
    }
    sync();
}

template <int LOG2N, int DONE, bool TAB, class Sync>
CORR_HD CORR_FI void stockham_rest(float2 (&v)[16], float2 *lds, const float2 *tw, const float2 *tab, int j,
                                   Sync &sync) {
    if constexpr (DONE < LOG2N) {
        constexpr int REM = LOG2N - DONE;
        constexpr int LOGR = REM >= 4 ? 4 : REM;
        stockham_pass<(1 << LOG2N), (1 << LOGR), (1 << DONE), TAB>(v, lds, tw, tab, j, sync);
        stockham_rest<LOG2N, DONE + LOGR, TAB>(v, lds, tw, tab, j, sync);
    }
}

// Validity bit of frame f of station st: bit (bit0 + f) of mask + st*mask_pitch,
// LSB first within a byte (reference kernels.cu:163-166).
CORR_HD CORR_FI int32_t mask_bit(const uint8_t *mask, size_t mask_pitch, int bit0, int f, int st) {
    const int bit = bit0 + f;
    return (int32_t)((mask[(size_t)st * mask_pitch + (size_t)(bit >> 3)] >> (bit & 7)) & 1u);
}

// Expand the validity bits of `frames` frames into one int32 per (station,
// frame): valid[st*valid_stride + f]. Thread f of the workgroup of station st.
CORR_HD CORR_FI void expand_validity(int f, int st, int frames, const uint8_t *mask, size_t mask_pitch,
                                    int bit0, int32_t *valid, size_t valid_stride) {
    if (f < frames) valid[(size_t)st * valid_stride + (size_t)f] = mask_bit(mask, mask_pitch, bit0, f, st);
}

// 4-bit staging (corrx_create_packed): a 32-bit word of packed samples holds
// samples 8i..8i+7, sample 2m in bits 0-3 and 2m+1 in bits 4-7 of byte m, two's
// complement. Returned as eight int8 bytes, sample 8i in the low byte: the
// staging layout of the same samples at 8 bits.
CORR_HD CORR_FI uint64_t unpack4_word(uint32_t w) {
    uint64_t out = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t n = (w >> (4 * k)) & 15u;
        const uint32_t b = (n ^ 8u) - 8u;  // the nibble sign-extended to 32 bits
        out |= (uint64_t)(b & 0xFFu) << (8 * k);
    }
    return out;
}

// Word i of station st's first `words` packed words (pitches in words; the
// packed and unpacked pitches hold the same number of samples).
CORR_HD CORR_FI void unpack4(size_t i, int st, const uint32_t *in, size_t in_pitch, uint64_t *out,
                            size_t out_pitch, size_t words) {
    if (i < words) out[(size_t)st * out_pitch + i] = unpack4_word(in[(size_t)st * in_pitch + i]);
}

// An int8 pair as two floats, exactly. On the device each half is one
// conversion that selects and sign-extends its byte of the 16-bit pair (SDWA
// BYTE_0 / BYTE_1); the compiler otherwise extracts and extends each byte
// with a shift or bit-field instruction first, two instructions per sample.
// ALU operations only: the pair is already in a register.
CORR_HD CORR_FI float2 int8_pair_to_float2(char2 c) {
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned w = (unsigned)(uint8_t)c.x | ((unsigned)(uint8_t)c.y << 8);
    float lo, hi;
    asm("v_cvt_f32_i32_sdwa %0, sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_0"
        : "=v"(lo)
        : "v"(w));
    asm("v_cvt_f32_i32_sdwa %0, sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_1"
        : "=v"(hi)
        : "v"(w));
    return make_float2(lo, hi);
#else
    return make_float2((float)c.x, (float)c.y);
#endif
}

// Thread j's 16 input values of frame f of station st, z[j + r*N/16] for
// r < 16, converted exactly from int8 pairs.
// Each load is the base of its run r (common to the workgroup) plus the lane's
// 32-bit byte offset 2j: scalar address arithmetic only.
template <int LOG2N>
CORR_HD CORR_FI void load_frame(int j, const int8_t *in, size_t station_pitch, int f, int st, float2 (&v)[16]) {
    constexpr int NT = (1 << LOG2N) / 16;
    const int8_t *src = in + (size_t)st * station_pitch + (size_t)f * (size_t)(2 << LOG2N);
    const unsigned oj = 2u * (unsigned)j;
    char2 raw[16];
#if defined(__HIP_DEVICE_COMPILE__)
    const __amdgpu_buffer_rsrc_t rs = raw_rsrc(src);
#pragma unroll
    for (int r = 0; r < 16; ++r)
        raw[r] = __builtin_bit_cast(char2, (unsigned short)__builtin_amdgcn_raw_buffer_load_b16(rs, oj, 2 * r * NT, 0));
#else
#pragma unroll
    for (int r = 0; r < 16; ++r) raw[r] = *reinterpret_cast<const char2 *>(src + 2 * r * NT + oj);
#endif
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = int8_pair_to_float2(raw[r]);
}

// Frame f of station st is present: `valid` is NULL (every frame present) or
// the expanded validity of expand_validity(), one aligned int32 per (station,
// frame), read once per workgroup at a uniform address.
CORR_HD CORR_FI bool frame_present(const int32_t *valid, size_t valid_stride, int f, int st) {
    return valid == nullptr || valid[(size_t)st * valid_stride + (size_t)f] != 0;
}

// The Z spectrum of thread j's converted input values into `lds` (natural
// order, padded); on return `lds` holds Z after a final barrier. With TAB the
// second pass's twiddles are read from `tab`, filled before the first barrier.
template <int LOG2N, bool TAB = false, class Sync>
CORR_HD CORR_FI void fft_from_values(int j, float2 (&v)[16], float2 *lds, Sync &sync, const float2 *tw,
                                     const float2 *tab = nullptr) {
    static_assert(LOG2N >= 9 && LOG2N <= 14, "the plain-layout R2C reads need a last pass with NS >= 256");
    // first pass: radix 16, NS = 1, inputs z[j + r*N/16] already in registers
    dft<16>(v);
#pragma unroll
    for (int r = 0; r < 16; ++r) lds[17 * j + r] = v[r];  // padded layout: pad16(16 j + r)
    sync();
    stockham_rest<LOG2N, 4, TAB>(v, lds, tw, tab, j, sync);
}

// The Z spectrum of frame f of station st into `lds` (natural order, padded).
// Thread j of N/16; on return `lds` holds Z after a final barrier. `tab`
// (kTw16Elems float2 past the exchange buffer) receives the second pass's
// twiddles.
template <int LOG2N, class Sync>
CORR_HD CORR_FI void complex_fft_block(int j, int f, int st, float2 *lds, Sync &sync, const int8_t *in,
                                      size_t station_pitch, const float2 *tw, float2 *tab) {
    float2 v[16];
    load_frame<LOG2N>(j, in, station_pitch, f, st, v);
    fill_tw16<(1 << LOG2N)>(j, tab, tw);
    fft_from_values<LOG2N, true>(j, v, lds, sync, tw, tab);
}

// R2C bin k from Z in lds (plain layout: the last pass of every supported
// length combines NS >= 256 points): (Z[k] + conj(Z[N-k]))/2 + W^k (Z[k] - conj(Z[N-k]))/(2i).
template <int N>
CORR_HD CORR_FI float2 r2c_bin(const float2 *lds, const float2 *wl, int k) {
    const float2 a = lds[k];
    const float2 b = lds[(N - k) & (N - 1)];
    const float er = 0.5f * (a.x + b.x), ei = 0.5f * (a.y - b.y);
    const float orr = 0.5f * (a.y + b.y), oi = -0.5f * (a.x - b.x);
    const float2 w = wl[(unsigned)k];
    return make_float2(er + (w.x * orr - w.y * oi), ei + (w.x * oi + w.y * orr));
}

// Bins k and N-k together (1 <= k < N): with E = (Z[k] + conj(Z[N-k]))/2 and
// O = (Z[k] - conj(Z[N-k]))/(2i), X[k] = E + W^k O and, as W^(N-k) = -conj(W^k),
// X[N-k] = conj(E - W^k O).
// The same from a = Z[k], b = Z[N-k] and wh = W^k / 2: the halvings of E ride
// in the final FMAs and those of O in wh, 12 instructions per pair where the
// textbook form takes 16. The roundings are the textbook form's: every
// halving is exact, so each product and sum rounds the same real value.
CORR_HD CORR_FI void r2c_combine(float2 a, float2 b, float2 wh, float2 &xk, float2 &xm) {
    const float s = a.x + b.x, d = a.y - b.y;    // 2 E
    const float orr = a.y + b.y, oi = b.x - a.x;  // 2 O
    const float tr = __builtin_fmaf(wh.x, orr, -(wh.y * oi));
    const float ti = __builtin_fmaf(wh.x, oi, wh.y * orr);
    xk = make_float2(__builtin_fmaf(0.5f, s, tr), __builtin_fmaf(0.5f, d, ti));
    xm = make_float2(__builtin_fmaf(0.5f, s, -tr), __builtin_fmaf(-0.5f, d, ti));
}

// W^k / 2 for k = j + q*N/16: wj / 2 for q = 0, else wj times the constant
// exp(-i pi q/16) / 2.
CORR_HD CORR_FI float2 half_twiddle(float2 wj, int q) {
    return q == 0 ? make_float2(0.5f * wj.x, 0.5f * wj.y)
                  : cmul(wj, make_float2(0.5f * cos32(q), -0.5f * cos32((q + 24) & 31)));
}

template <int N>
CORR_HD CORR_FI void r2c_pair(const float2 *lds, float2 wh, int k, float2 &xk, float2 &xm) {
    r2c_combine(lds[k], lds[N - k], wh, xk, xm);
}

// The R2C epilogue of one workgroup: every retained bin k = 0..N-1 exactly once,
// thread j taking k = j + q*N/16 (q < 8) with its mirror N - k, thread 0 also
// the self-mirrored bin N/2 (bin 0's mirror is the Nyquist bin, never formed).
// W^k = W^j W^(q N/16) with W^(q N/16) = exp(-i pi q/16), a constant: one
// table read per thread. The emitter stores them: emit.pair(q, X[k], X[N-k])
// for k = j + q*N/16 != 0, emit.at(k, X[k]) for bins 0 and N/2.
template <int N, class Emit>
CORR_HD CORR_FI void r2c_epilogue(int j, const float2 *lds, const float2 *wl, Emit &&emit) {
    constexpr int NT = N / 16;
    const float2 wj = wl[(unsigned)j];
    emit.begin();
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int k = j + q * NT;
        if (k == 0) {
            emit.at(0, r2c_bin<N>(lds, wl, 0));
        } else {
            // exp(-i pi q/16) = cos(2 pi q/32) - i sin(2 pi q/32)
            float2 xk, xm;
            r2c_pair<N>(lds, half_twiddle(wj, q), k, xk, xm);
            emit.pair(q, xk, xm);
        }
    }
    if (j == 0) emit.at(N / 2, r2c_bin<N>(lds, wl, N / 2));
}

// The radix-16 Stockham passes from DONE points combined while 16 or more
// remain; each ends with a barrier.
template <int LOG2N, int DONE, bool TAB, class Sync>
CORR_HD CORR_FI void stockham_radix16_passes(float2 (&v)[16], float2 *lds, const float2 *tw, const float2 *tab, int j,
                                             Sync &sync) {
    if constexpr (LOG2N - DONE >= 4) {
        stockham_pass<(1 << LOG2N), 16, (1 << DONE), TAB>(v, lds, tw, tab, j, sync);
        stockham_radix16_passes<LOG2N, DONE + 4, TAB>(v, lds, tw, tab, j, sync);
    }
}

// The last pass (radix R = 2^(LOG2N mod 4) = 2, 4 or 8, NS = N/R) and the R2C
// unpack, in registers. Butterfly b makes Z[b + NS r], r < R, and the mirrors
// N - k of those bins are butterfly NS - b's outputs with r -> R - 1 - r. Lane
// j takes the NB/2 butterflies b_q = j + (N/16) q and their mirrors NS - b_q
// (lane 0: NS/2 for b_0 = 0, the one butterfly no lane pairs), so both bins of
// each of its R2C pairs are in its registers. The mirrors' twiddles are
// W^(NS r) = exp(-2 pi i r/R) times the conjugates of the direct ones. `lds`
// holds the previous pass's outputs after its barrier; nothing is written back.
template <int LOG2N, class Emit>
CORR_HD CORR_FI void last_pass_r2c(int j, const float2 *lds, const float2 *tw, const float2 *wl, Emit &&emit) {
    constexpr int N = 1 << LOG2N, NT = N / 16;
    constexpr int LOGR = LOG2N % 4, R = 1 << LOGR, NB = 16 / R, H = NB / 2, NS = N / R;
    constexpr bool PIN = pad_for(NS / 16);
    static_assert(LOGR >= 1, "a last pass of radix 2, 4 or 8");
    float2 v[16];
    {
        const float2 *src = lds + lds_index<PIN>(j);
        const float2 *srm = lds + lds_index<PIN>(j == 0 ? NS / 2 : NS - j);  // mirror of b_0
        const float2 *srn = lds + lds_index<PIN>(NS - j);                       // b_q mirrors, q >= 1
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int q = 0; q < H; ++q) v[q * R + r] = src[lds_span<PIN>(q * NT + r * NS)];
            v[H * R + r] = srm[lds_span<PIN>(r * NS)];
#pragma unroll
            for (int q = 1; q < H; ++q) v[(H + q) * R + r] = srn[lds_span<PIN>(r * NS) - lds_span<PIN>(q * NT)];
        }
    }
    {
        float2 w0[R - 1];  // W^(j r), W = exp(-2 pi i/N)
        group_twiddles<R, 1>(tw, j, w0);
#pragma unroll
        for (int q = 0; q < H; ++q) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const int m = (2 * r * q) & 31;  // W^(N/16 q r) = exp(-2 pi i m/32)
                const float2 wd = q == 0 ? w0[r - 1] : cmul(w0[r - 1], make_float2(cos32(m), -cos32((m + 24) & 31)));
                float2 wm;
                const float2 c = make_float2(wd.x, -wd.y);
                switch ((32 / R) * r) {  // W^(NS r) conj(wd), NS r/N = r/R turns
                    case 2: wm = rot32_any<2>(c); break;
                    case 4: wm = rot32_any<4>(c); break;
                    case 6: wm = rot32_any<6>(c); break;
                    case 8: wm = rot32_any<8>(c); break;
                    case 10: wm = rot32_any<10>(c); break;
                    case 12: wm = rot32_any<12>(c); break;
                    case 14: wm = rot32_any<14>(c); break;
                    case 16: wm = rot32_any<16>(c); break;
                    case 20: wm = rot32_any<20>(c); break;
                    case 24: wm = rot32_any<24>(c); break;
                    case 28: wm = rot32_any<28>(c); break;
                    default: wm = c; break;
                }
                if (q == 0 && j == 0) {
                    // butterfly NS/2: W^(NS/2 r) = exp(-i pi r/R)
                    const int mh = (16 / R) * r;
                    wm = make_float2(cos32(mh), -cos32((mh + 24) & 31));
                }
                v[q * R + r] = cmul(v[q * R + r], wd);
                v[(H + q) * R + r] = cmul(v[(H + q) * R + r], wm);
            }
        }
#pragma unroll
        for (int g = 0; g < NB; ++g) dft<R>(&v[g * R]);
    }
    // v[q R + r] = Z[b_q + NS r], v[(H + q) R + r] = Z[N - b_q - NS (R - 1 - r)]; the pair of
    // k = b_q + NS r = j + (N/16)(q + NB r) is emitted as q' = q + NB r.
    const float2 wj = wl[(unsigned)j];
    emit.begin();
    // the butterflies every lane has alike first, so that their stores' offsets
    // are formed before the branch on the bin-0 lane, not on both of its sides
#pragma unroll
    for (int q = 1; q < H; ++q)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float2 xk, xm;
            r2c_combine(v[q * R + r], v[(H + q) * R + R - 1 - r], half_twiddle(wj, q + NB * r), xk, xm);
            emit.pair(q + NB * r, xk, xm);
        }
    if (j != 0) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float2 xk, xm;
            r2c_combine(v[r], v[H * R + R - 1 - r], half_twiddle(wj, NB * r), xk, xm);
            emit.pair(NB * r, xk, xm);
        }
    } else {
        // the pairs (NS r, N - NS r) inside butterfly 0 and (NS/2 + NS r, its
        // mirror) inside butterfly NS/2, then bin 0 (X = Re Z[0] + Im Z[0]) and
        // bin N/2 (X = conj Z[N/2], Z[N/2] = v[R/2]). Lane 0 stores them with
        // plain stores only: the branches then share no store the compiler
        // could merge after them into one whose base differs between lanes.
#pragma unroll
        for (int r = 1; r < R / 2; ++r) {
            float2 xk, xm;
            r2c_combine(v[r], v[R - r], half_twiddle(wj, NB * r), xk, xm);
            emit.at(NB * r * NT, xk);
            emit.at(N - NB * r * NT, xm);
        }
#pragma unroll
        for (int r = 0; r < R / 2; ++r) {
            float2 xk, xm;
            r2c_combine(v[H * R + r], v[H * R + R - 1 - r], half_twiddle(wj, H + NB * r), xk, xm);
            emit.at((H + NB * r) * NT, xk);
            emit.at(N - (H + NB * r) * NT, xm);
        }
        emit.at(0, make_float2(v[0].x + v[0].y, 0.0f));
        emit.at(N / 2, make_float2(v[R / 2].x, -v[R / 2].y));
    }
}

// Whole F-engine workgroup, complex64 output [station][frame][N].
template <int LOG2N, class Sync, class Emit>
CORR_HD CORR_FI void fengine_frame(int j, int f, int st, float2 *lds, Sync &sync, const int8_t *in,
                                   size_t station_pitch, const int32_t *valid, size_t valid_stride,
                                   const float2 *tw, const float2 *wl, Emit &&emit);

// Emitters store the R2C bins of lane j: begin() before the first store,
// at(k, X) bin k (bins 0 and N/2 only), pair(q, X[k], X[N-k]) for
// k = j + q*N/16; zero_at / zero_pair store zero bins (a missing frame). Each bin's address is a base that
// depends on q alone, the same for the whole workgroup, plus the lane's own
// fixed offset: the base is scalar arithmetic, so a store costs no vector
// address arithmetic. The mirror bin N - j - q*N/16 is addressed as
// base(N - (q+1)*N/16) + offset(N/16 - j).
template <int LOG2N, int FORM, class Sync>
CORR_HD CORR_FI void fengine_form_block(int j, int f, int st, float2 *lds, Sync &sync, const int8_t *in,
                                       size_t station_pitch, const int32_t *valid, size_t valid_stride,
                                       const float2 *tw, const float2 *wl, const float *gain,
                                       void *out, size_t out_station_stride, int tile_shift);

// Whole F-engine workgroup, complex64 output: [station][frame][N] (tile_shift =
// LOG2N, the plain order), or tiles of 2^tile_shift channels in the order of
// forms::spec_index, out_station_stride = N * frame capacity.
template <int LOG2N, class Sync>
CORR_HD CORR_FI void fengine_fp32_block(int j, int f, int st, float2 *lds, Sync &sync, const int8_t *in,
                                       size_t station_pitch, const int32_t *valid, size_t valid_stride,
                                       const float2 *tw, const float2 *wl, float2 *out,
                                       size_t out_station_stride, int tile_shift = LOG2N) {
    fengine_form_block<LOG2N, 0>(j, f, st, lds, sync, in, station_pitch, valid, valid_stride, tw, wl, nullptr, out,
                                 out_station_stride, tile_shift);
}

// The form's emitter: bin k scaled by its gain and converted (forms.hpp), at
// offset off(k) = (k >> shift) * tile_stride + k % 2^shift from the frame's
// base. off() is additive over a multiple of the tile plus a remainder, which
// the pair addressing needs: the tile is at most N/16 channels, or N.
// FORM 0 is the FP32 form: the bin itself as complex64, no gain.
template <int N, int FORM>
struct FormEmit {
    using word_t = typename std::conditional<
        FORM == 0, float2, typename std::conditional<FORM == 1, uint32_t, uint16_t>::type>::type;
    word_t *dst;         // the frame's base in the station's tile-major block
    const float *g;      // the station's gains (not read in the FP32 form)
    unsigned shift, mask, tile_stride;
    // Byte offsets of the lane's bins j and N/16 - j from a base common to the
    // workgroup: the spectrum words off(j), off(N/16 - j) and the gains j,
    // N/16 - j. They fit 32 bits (off(N/16) * 8 = N * frame capacity / 2 bytes,
    // below 2^31 as N * frame capacity < 2^32), so every access in pair() is
    // a scalar base plus the lane's 32-bit offset: no vector address arithmetic.
    unsigned ok = 0, om = 0, gk = 0, gm = 0;
    int j;
    CORR_HD CORR_FI unsigned off(unsigned k) const { return (k >> shift) * tile_stride + (k & mask); }
    // The lane offsets are formed here, after the transform, so that they do
    // not hold registers through it.
    CORR_HD CORR_FI void begin() {
        ok = off((unsigned)j) * (unsigned)sizeof(word_t);
        om = off((unsigned)(N / 16 - j)) * (unsigned)sizeof(word_t);
        gk = (unsigned)j * 4u;
        gm = (unsigned)(N / 16 - j) * 4u;
    }
    // An element index that is the same in every lane (a function of the run's
    // tile parameters and a compile-time bin), read from the first active lane.
    // That states its uniformity to the compiler where it cannot prove it:
    // after the branches on the bin-0 lane it may merge the indices of the two
    // sides, and a buffer access whose base it then takes for a per-lane value
    // becomes a loop over the wave's distinct bases.
    CORR_HD CORR_FI static unsigned uniform(unsigned x) {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_amdgcn_readfirstlane(x);
#else
        return x;
#endif
    }
    // The gain at byte b from base (1 in the FP32 form, which reads none).
    CORR_HD CORR_FI static float gain(const float *base, unsigned b) {
        if constexpr (FORM == 0) {
            return 1.0f;
        } else {
#if defined(__HIP_DEVICE_COMPILE__)
            return __builtin_bit_cast(float, (unsigned)__builtin_amdgcn_raw_buffer_load_b32(raw_rsrc(base), b, 0, 0));
#else
            return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + b);
#endif
        }
    }
    // Store w at byte b from base.
    CORR_HD CORR_FI static void put(word_t *base, unsigned b, word_t w) {
#if defined(__HIP_DEVICE_COMPILE__)
        typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
        if constexpr (FORM == 0)
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, w), raw_rsrc(base), b, 0, 0);
        else if constexpr (FORM == 1)
            __builtin_amdgcn_raw_buffer_store_b32(w, raw_rsrc(base), b, 0, 0);
        else
            __builtin_amdgcn_raw_buffer_store_b16(w, raw_rsrc(base), b, 0, 0);
#else
        *reinterpret_cast<word_t *>(reinterpret_cast<char *>(base) + b) = w;
#endif
    }
    // The bin's word with gain gn: the gain is read only by the requantizing forms
    CORR_HD CORR_FI static word_t word_g(float2 x, float gn) {
        if constexpr (FORM == 0)
            return x;
        else if constexpr (FORM == 1)
            return forms::pack_fp16(x, gn);
        else
            return forms::pack_int8(x, gn);
    }
    // bin k's word (gain index k)
    CORR_HD CORR_FI word_t word(float2 x, unsigned k) const {
        if constexpr (FORM == 0)
            return x;
        else
            return word_g(x, g[k]);
    }
    CORR_HD CORR_FI static word_t zero() {
        if constexpr (FORM == 0)
            return make_float2(0.0f, 0.0f);
        else
            return (word_t)0;  // zero requantizes to the all-zero word
    }
    CORR_HD CORR_FI void at(int k, float2 x) const { dst[off((unsigned)k)] = word(x, (unsigned)k); }
    CORR_HD CORR_FI void pair(int q, float2 xk, float2 xm) const {
        constexpr int NT = N / 16;
        const unsigned bk = (unsigned)(q * NT), bm = (unsigned)(N - (q + 1) * NT);
        put(dst + uniform(off(bk)), ok, word_g(xk, gain(g + bk, gk)));
        put(dst + uniform(off(bm)), om, word_g(xm, gain(g + bm, gm)));
    }
    CORR_HD CORR_FI void zero_at(int k) const { dst[off((unsigned)k)] = zero(); }
    CORR_HD CORR_FI void zero_pair(int q) const {
        constexpr int NT = N / 16;
        put(dst + uniform(off((unsigned)(q * NT))), ok, zero());
        put(dst + uniform(off((unsigned)(N - (q + 1) * NT))), om, zero());
    }
};

// Whole F-engine workgroup for the FP16 (FORM 1) or int8 (FORM 2) form: bin k
// of station st is scaled by gain[st*N + k] and converted (forms.hpp); output
// packed 32-bit (FP16) or 16-bit (int8) complex in the tile-major order of
// forms::spec_index with tiles of 2^tile_shift channels, out_station_stride =
// N * frame capacity.
// FORM 0: the FP32 form's complex64 bins in the same tile-major order (gain unused).
template <int LOG2N, int FORM, class Sync>
CORR_HD CORR_FI void fengine_form_block(int j, int f, int st, float2 *lds, Sync &sync, const int8_t *in,
                                       size_t station_pitch, const int32_t *valid, size_t valid_stride,
                                       const float2 *tw, const float2 *wl, const float *gain,
                                       void *out, size_t out_station_stride, int tile_shift) {
    constexpr int N = 1 << LOG2N;
    using Emit = FormEmit<N, FORM>;
    // bin k -> tile k >> shift at (tile * fcap + f) * T + k % T from the station's base
    Emit e;
    e.dst = static_cast<typename Emit::word_t *>(out) + (size_t)st * out_station_stride +
            ((size_t)f << tile_shift);
    e.g = FORM == 0 ? gain : gain + (size_t)st * N;
    e.shift = (unsigned)tile_shift;
    e.mask = (1u << tile_shift) - 1u;
    e.tile_stride = (unsigned)(out_station_stride / N) << tile_shift;  // fcap * T
    e.j = j;
    fengine_frame<LOG2N>(j, f, st, lds, sync, in, station_pitch, valid, valid_stride, tw, wl, e);
}

// ---------------------------------------------------------------------------
// Half-buffer exchanges. At N = 16384 a full exchange buffer (136 KB) leaves
// room for one workgroup per CU, whose barriers and first HBM read then idle
// the CU. Here every exchange goes through a buffer of N/2 points in two rounds
// (write the values bound for one half, barrier, read the values taken from
// it), so two workgroups share a CU and each hides the other's stalls. Which
// half a value belongs to is fixed by the lane's wave or by the radix digit,
// never data-dependent, so no wave diverges.
// ---------------------------------------------------------------------------
template <int N>
struct HalfSizes {
    static constexpr int kLds = N / 2 + N / 32;  // padded half buffer, float2
};

// Write the outputs of pass (R, NS) held in v that fall in half h of [0, N).
// R = 16 (one group per lane): the half is set by the lane; the last pass
// (NS*R = N): by the digit r.
template <int N, int R, int NS>
CORR_HD CORR_FI void half_write(const float2 (&v)[16], float2 *lds, int j, int h) {
    constexpr int NT = N / 16;
    constexpr int NB = 16 / R;
    constexpr bool POUT = pad_for(NS);
    if constexpr (R == 16) {
        const int k = j & (NS - 1);
        const int base = (j - k) * 16 + k;  // < N/2 exactly when the lane's writes are
        if ((base >= N / 2) == (h == 1)) {
            int at = lds_index<POUT>(base - h * (N / 2));
#if defined(__HIP_DEVICE_COMPILE__)
            // a base of its own for each half: derived from the other half's,
            // the second half's stores would take offsets below it (negative,
            // which a ds_write's immediate cannot hold) and one address add each
            asm volatile("" : "+v"(at));
#endif
            float2 *dst = lds + at;
#pragma unroll
            for (int r = 0; r < 16; ++r) dst[NS == 1 ? r : lds_span<POUT>(r * NS)] = v[r];
        }
    } else {
        static_assert(NS * R == N, "a pass of radix below 16 is the last pass");
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            float2 *dst = lds + lds_index<POUT>(j + q * NT);
#pragma unroll
            for (int r = 0; r < R; ++r)
                if ((2 * r >= R) == (h == 1)) dst[lds_span<POUT>(r * NS - h * (N / 2))] = v[q * R + r];
        }
    }
}

// Read the inputs of the pass of radix R (stride N/R) that lie in half h.
template <int N, int R, bool PIN>
CORR_HD CORR_FI void half_read(float2 (&v)[16], const float2 *lds, int j, int h) {
    constexpr int NT = N / 16;
    constexpr int NB = 16 / R;
    constexpr int STRIDE = N / R;
    const float2 *src = lds + lds_index<PIN>(j);
#pragma unroll
    for (int q = 0; q < NB; ++q)
#pragma unroll
        for (int r = 0; r < R; ++r)
            if ((2 * r >= R) == (h == 1)) v[q * R + r] = src[lds_span<PIN>(q * NT + r * STRIDE - h * (N / 2))];
}

// One exchange from pass (R0, NS0)'s outputs to the inputs of the next pass of
// radix R1, through the half buffer. `first` skips the barrier that protects
// the previous exchange's reads.
template <int N, int R0, int NS0, int R1, class Sync>
CORR_HD CORR_FI void half_exchange(float2 (&v)[16], float2 *lds, int j, Sync &sync, bool first) {
    float2 in[16];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (h == 1 || !first) sync();
        half_write<N, R0, NS0>(v, lds, j, h);
        sync();
        half_read<N, R1, pad_for(NS0)>(in, lds, j, h);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = in[r];
}

// Pass (R, NS) arithmetic on inputs already in registers: twiddles, DFT.
template <int N, int R, int NS>
CORR_HD CORR_FI void pass_compute(float2 (&v)[16], const float2 (&w)[16 / R][R - 1]) {
    constexpr int NB = 16 / R;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[q * R + r] = cmul(v[q * R + r], w[q][r - 1]);
        dft<R>(&v[q * R]);
    }
}

// A pass's twiddles in two steps: the table reads (issued before the
// exchange, so that they are in flight during it) and the products (after it).
template <int N, int R, int NS>
CORR_HD CORR_FI void pass_twiddle_reads(const float2 *tw, int j, float2 (&w)[16 / R][R - 1]) {
    constexpr int NT = N / 16;
    constexpr int NB = 16 / R;
    constexpr int TWSTEP = N / (NS * R);
    if constexpr (NS * R == N && NB > 1) {
        group_twiddle_reads<R, TWSTEP>(tw, j, w[0]);
    } else {
#pragma unroll
        for (int q = 0; q < NB; ++q) group_twiddle_reads<R, TWSTEP>(tw, (j + q * NT) & (NS - 1), w[q]);
    }
}

template <int N, int R, int NS>
CORR_HD CORR_FI void pass_twiddle_products(float2 (&w)[16 / R][R - 1]) {
    constexpr int NB = 16 / R;
    if constexpr (NS * R == N && NB > 1) {
        group_twiddle_products<R>(w[0]);
#pragma unroll
        for (int q = 1; q < NB; ++q)
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const int m = (2 * r * q) & 31;
                w[q][r - 1] = cmul(w[0][r - 1], make_float2(cos32(m), -cos32((m + 24) & 31)));
            }
    } else {
#pragma unroll
        for (int q = 0; q < NB; ++q) group_twiddle_products<R>(w[q]);
    }
}

// The N = 16384 transform after its first pass (v: the second pass's inputs,
// z'[j + r*N/16], already exchanged): passes 16, 16, 4 and the R2C epilogue,
// all exchanges through the half buffer `lds`. `tab` holds the second pass's
// twiddles (fill_tw16, before the first barrier). emit(k, X) stores one bin.
template <class Sync, class Emit>
CORR_HD CORR_FI void fengine_half_16384_rest(int j, float2 (&v)[16], float2 *lds, const float2 *tab, Sync &sync,
                                             const float2 *tw, const float2 *wl, Emit &&emit) {
    constexpr int N = 16384;
    {
        float2 w[1][15];
        read_tw16(tab, j, w[0]);
        pass_compute<N, 16, 16>(v, w);
    }
    {
        float2 w[1][15];
        pass_twiddle_reads<N, 16, 256>(tw, j, w);
        half_exchange<N, 16, 16, 16>(v, lds, j, sync, false);
        pass_twiddle_products<N, 16, 256>(w);
        pass_compute<N, 16, 256>(v, w);
    }
    // Last pass (radix 4, NS = 4096) and the R2C unpack in registers. Butterfly
    // b of the last pass makes Z[b + 4096 r], r < 4, and the mirrors N - k of
    // those bins are the outputs of butterfly 4096 - b (r -> 3 - r). Lane j takes
    // the butterflies b0 = j, b1 = 4096 - j, b2 = 1024 + j, b3 = 3072 - j (lane 0:
    // b1 = 2048, the one butterfly no lane pairs), so it holds both bins of each
    // of its R2C pairs and the spectrum never goes back through LDS.
    const int b1 = j == 0 ? 2048 : 4096 - j;
    float2 w0[3];  // W^(j r), r = 1..3 (W = exp(-2 pi i/N)): table w, w^2 and their product
    w0[0] = tw[(unsigned)j];
    w0[1] = tw[(unsigned)(2 * j)];
    {
        float2 in[16];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            sync();
            half_write<N, 16, 256>(v, lds, j, h);  // the third pass's outputs, plain layout
            sync();
#pragma unroll
            for (int r = 2 * h; r < 2 * h + 2; ++r) {
                const int o = r * 4096 - h * (N / 2);
                in[r] = lds[j + o];
                in[4 + r] = lds[b1 + o];
                in[8 + r] = lds[1024 + j + o];
                in[12 + r] = lds[3072 - j + o];
            }
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = in[r];
    }
    {
        w0[2] = cmul(w0[0], w0[1]);
        // b2 = 1024 + j: W^(1024 r) W^(j r), W^(1024 r) = exp(-i pi r/8)
        float2 w2[3];
#pragma unroll
        for (int r = 1; r < 4; ++r) w2[r - 1] = cmul(w0[r - 1], make_float2(cos32(2 * r), -cos32((2 * r + 24) & 31)));
        // b1 = 4096 - j and b3 = 3072 - j: W^(4096 r) = (-i)^r times the conjugate
        // of b0's and b2's twiddles; lane 0's b1 = 2048: W^(2048 r) = exp(-i pi r/4)
        float2 w1[3], w3[3];
#pragma unroll
        for (int r = 1; r < 4; ++r) {
            const float2 c0 = w0[r - 1], c2 = w2[r - 1];
            // (-i)^r conj(c): r = 1: (-c.y, -c.x); r = 2: (-c.x, c.y); r = 3: (c.y, c.x)
            w1[r - 1] = r == 1 ? make_float2(-c0.y, -c0.x) : r == 2 ? make_float2(-c0.x, c0.y) : make_float2(c0.y, c0.x);
            w3[r - 1] = r == 1 ? make_float2(-c2.y, -c2.x) : r == 2 ? make_float2(-c2.x, c2.y) : make_float2(c2.y, c2.x);
            if (j == 0) w1[r - 1] = make_float2(cos32(4 * r), -cos32((4 * r + 24) & 31));
        }
#pragma unroll
        for (int r = 1; r < 4; ++r) {
            v[r] = cmul(v[r], w0[r - 1]);
            v[4 + r] = cmul(v[4 + r], w1[r - 1]);
            v[8 + r] = cmul(v[8 + r], w2[r - 1]);
            v[12 + r] = cmul(v[12 + r], w3[r - 1]);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) dft<4>(&v[4 * q]);
    }
    // v[4q + r] = Z[b_q + 4096 r]. The pairs (k, N - k): k = j + 4096 r with
    // N - k = b1 + 4096 (3 - r), and k = 1024 + j + 4096 r with N - k =
    // b3 + 4096 (3 - r); k = j + 1024 q' for q' = 4r and 4r + 1, so the emitter
    // and W^k = W^j exp(-i pi q'/16) follow q' (W = exp(-i pi/N)). Lane 0 has
    // bin 0 (X = Re Z[0] + Im Z[0]), bin N/2 (X = conj Z[N/2]) and the pairs
    // (4096, 12288), (2048, 14336), (6144, 10240) in place of its first four.
    const float2 wj = wl[(unsigned)j];
    emit.begin();
    if (j != 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float2 xk, xm;
            r2c_combine(v[r], v[4 + 3 - r], half_twiddle(wj, 4 * r), xk, xm);
            emit.pair(4 * r, xk, xm);
        }
    } else {
        // plain stores only, as in last_pass_r2c
        float2 xk, xm;
        r2c_combine(v[1], v[3], half_twiddle(wj, 4), xk, xm);
        emit.at(4096, xk);
        emit.at(N - 4096, xm);
        r2c_combine(v[4], v[7], half_twiddle(wj, 2), xk, xm);
        emit.at(2048, xk);
        emit.at(N - 2048, xm);
        r2c_combine(v[5], v[6], half_twiddle(wj, 6), xk, xm);
        emit.at(6144, xk);
        emit.at(N - 6144, xm);
        emit.at(0, make_float2(v[0].x + v[0].y, 0.0f));
        emit.at(N / 2, make_float2(v[2].x, -v[2].y));
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        float2 xk, xm;
        r2c_combine(v[8 + r], v[12 + 3 - r], half_twiddle(wj, 4 * r + 1), xk, xm);
        emit.pair(4 * r + 1, xk, xm);
    }
}

// The whole N = 16384 transform with the first pass on the vector ALUs: v holds
// lane j's inputs z[j + r*N/16].
template <class Sync, class Emit>
CORR_HD CORR_FI void fengine_half_16384(int j, float2 (&v)[16], float2 *lds, const float2 *tab, Sync &sync,
                                        const float2 *tw, const float2 *wl, Emit &&emit) {
    dft<16>(v);  // first pass: radix 16, NS = 1
    half_exchange<16384, 16, 1, 16>(v, lds, j, sync, true);
    fengine_half_16384_rest(j, v, lds, tab, sync, tw, wl, emit);
}


// LDS elements (float2) of the F-engine workgroup at N = 2^LOG2N.
// The power-of-two F-engine's LDS: the second pass's twiddle table
// (kTw16Elems float2, first, so that its reads take small immediate offsets
// from one address), then the exchange buffer.
template <int LOG2N>
CORR_HD constexpr int fengine_exchange_elems() {
    return LOG2N == 14 ? HalfSizes<(1 << LOG2N)>::kLds : Sizes<(1 << LOG2N)>::kLds;
}

template <int LOG2N>
CORR_HD constexpr int fengine_lds_elems() {
    return kTw16Elems + fengine_exchange_elems<LOG2N>();
}

// Lane j's bins of a missing frame, whose spectrum is zero, in the order the
// R2C epilogue emits them.
template <int N, class Emit>
CORR_HD CORR_FI void emit_zero_frame(int j, Emit &&emit) {
    emit.begin();
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        if (j + q * (N / 16) == 0)
            emit.zero_at(0);
        else
            emit.zero_pair(q);
    }
    if (j == 0) emit.zero_at(N / 2);
}

// One frame of one station through the F-engine: N = 16384 exchanges through
// the half buffer, shorter transforms through the full one. A missing frame is
// neither read nor transformed: its spectrum is written as zeros, the value the
// transform of a zeroed frame has (up to the sign of a zero, which changes no
// sum the X-engine forms). Presence is the same for the whole workgroup, so
// the early return splits no barrier.
template <int LOG2N, class Sync, class Emit>
CORR_HD CORR_FI void fengine_frame(int j, int f, int st, float2 *lds, Sync &sync, const int8_t *in,
                                   size_t station_pitch, const int32_t *valid, size_t valid_stride,
                                   const float2 *tw, const float2 *wl, Emit &&emit) {
    constexpr int N = 1 << LOG2N;
    if (!frame_present(valid, valid_stride, f, st)) {
        emit_zero_frame<N>(j, emit);
        return;
    }
    float2 *tab = lds, *xbuf = lds + kTw16Elems;
    if constexpr (LOG2N == 14) {
        float2 v[16];
        load_frame<LOG2N>(j, in, station_pitch, f, st, v);
        fill_tw16<N>(j, tab, tw);
        fengine_half_16384(j, v, xbuf, tab, sync, tw, wl, emit);
    } else {
        if constexpr (LOG2N == 13) {
            // first pass, the radix-16 passes, then the last pass and the unpack in
            // registers (at N = 8192; at N <= 2048, where two to four workgroups share a
            // CU, it measured slower than the unpack through LDS: more VALU work and
            // no exchange worth saving)
            float2 v[16];
            load_frame<LOG2N>(j, in, station_pitch, f, st, v);
            fill_tw16<N>(j, tab, tw);
            dft<16>(v);
#pragma unroll
            for (int r = 0; r < 16; ++r) xbuf[17 * j + r] = v[r];  // padded layout: pad16(16 j + r)
            sync();
            stockham_radix16_passes<LOG2N, 4, true>(v, xbuf, tw, tab, j, sync);
            last_pass_r2c<LOG2N>(j, xbuf, tw, wl, emit);
        } else {
            complex_fft_block<LOG2N>(j, f, st, xbuf, sync, in, station_pitch, tw, tab);
            r2c_epilogue<N>(j, xbuf, wl, emit);
        }
    }
}

// ---------------------------------------------------------------------------
// n_freq = M * 2^a with an odd factor M in {3, 5, 7}: the lengths 6144, 10240,
// 12288 and 14336 that the reference accepts (every multiple of 2048,
// corr-cuda/src/corr_cuda.cpp:500-509) besides the powers of two.
//
// The N = M*P point complex transform (P = 2^LOG2P) takes one radix-M stage
// by decimation in frequency, then M independent P-point Stockham transforms
// (the passes above), all in LDS:
//   n = n1 + P n2, k = M k1 + k2 (n1, k1 < P; n2, k2 < M)
//   y_k2[n1] = W_N^(n1 k2) sum_n2 z[n1 + P n2] W_M^(n2 k2)
//   Z[M k1 + k2] = sum_n1 y_k2[n1] W_P^(n1 k1)
// Sub-transform k2 lives in its own LDS span and runs on lanes
// [k2 P/16, (k2+1) P/16), whole waves, with the passes' own lane numbering.
// The R2C epilogue then reads Z[k] at (k mod M, k div M).
// ---------------------------------------------------------------------------
template <int M, int LOG2P>
struct MixedSizes {
    static_assert(M == 3 || M == 5 || M == 7, "odd factor 3, 5 or 7");
    static constexpr int P = 1 << LOG2P;
    static constexpr int N = M * P;
    static constexpr int kThreads = N / 16;
    static constexpr int kSubThreads = P / 16;
    // One sub-transform's span: the P-point exchange buffer plus 8 float2, so
    // that equal offsets in neighbouring spans fall in different LDS banks.
    static constexpr int kSub = Sizes<P>::kLds + 8;
    static constexpr int kLds = M * kSub;  // float2 elements
};

// cos(2 pi m/M) and sin(2 pi m/M) for 0 < m <= (M-1)/2, to 20 digits.
template <int M>
CORR_HD constexpr float cos_m(int m) {
    return M == 3   ? -0.5f
           : M == 5 ? (m == 1 ? 0.30901699437494742410f : -0.80901699437494742410f)
                    : (m == 1   ? 0.62348980185873353053f
                       : m == 2 ? -0.22252093395631440429f
                                : -0.90096886790241912624f);
}
template <int M>
CORR_HD constexpr float sin_m(int m) {
    return M == 3   ? 0.86602540378443864676f
           : M == 5 ? (m == 1 ? 0.95105651629515357212f : 0.58778525229247312917f)
                    : (m == 1   ? 0.78183148246802980871f
                       : m == 2 ? 0.97492791218182360702f
                                : 0.43388373911755812048f);
}

// In-place DFT of odd length M, natural order: with t_n = a_n + a_(M-n) and
// u_n = a_n - a_(M-n) (0 < n <= H = (M-1)/2),
//   X_0 = a_0 + sum t_n,
//   X_k = C_k - i S_k, X_(M-k) = C_k + i S_k (0 < k <= H),
//   C_k = a_0 + sum_n t_n cos(2 pi nk/M), S_k = sum_n u_n sin(2 pi nk/M),
// with nk reduced mod M into the first half by cos symmetry and sin antisymmetry.
template <int M>
CORR_HD CORR_FI void dft_odd(float2 (&a)[M]) {
    constexpr int H = (M - 1) / 2;
    float2 t[H + 1], u[H + 1];
#pragma unroll
    for (int n = 1; n <= H; ++n) {
        t[n] = cadd(a[n], a[M - n]);
        u[n] = csub(a[n], a[M - n]);
    }
    float2 x0 = a[0];
#pragma unroll
    for (int n = 1; n <= H; ++n) x0 = cadd(x0, t[n]);
    float2 out[M];
    out[0] = x0;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        float2 c = a[0], s = make_float2(0.0f, 0.0f);
#pragma unroll
        for (int n = 1; n <= H; ++n) {
            const int r = (n * k) % M;
            const int m = r <= H ? r : M - r;
            const float cs = cos_m<M>(m), sn = r <= H ? sin_m<M>(m) : -sin_m<M>(m);
            c = make_float2(c.x + t[n].x * cs, c.y + t[n].y * cs);
            s = make_float2(s.x + u[n].x * sn, s.y + u[n].y * sn);
        }
        out[k] = make_float2(c.x + s.y, c.y - s.x);
        out[M - k] = make_float2(c.x - s.y, c.y + s.x);
    }
#pragma unroll
    for (int k = 0; k < M; ++k) a[k] = out[k];
}

// The radix-M stage: lane j takes n1 = j + i*N/16 (< P), reads the M inputs
// z[n1 + P n2] of the frame (int8 pairs, exact in float), and writes
// y_k2[n1] = DFT_M(...)[k2] * W_N^(n1 k2) at offset n1 of span k2 (plain
// layout). twn[m] = exp(-2 pi i m/N), m < N; n1 k2 < N.
// Every load of the lane's NI columns is issued before the first is used, so
// that their latencies overlap (a column at a time, each waits for its own),
// and a column reads one twiddle, W_N^n1, whose powers W_N^(n1 k2), k2 < M,
// are formed by products (M - 2 complex products, each adding a rounding of
// about 2^-24): M - 1 scattered table reads per column took longer.
template <int M, int LOG2P>
CORR_HD CORR_FI void mixed_first_stage(int j, const char2 *src, float2 *lds, const float2 *twn) {
    using S = MixedSizes<M, LOG2P>;
    constexpr int NI = (S::P + S::kThreads - 1) / S::kThreads;
    char2 c[NI][M];
    float2 w[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int n1 = j + i * S::kThreads;
        if (i + 1 < NI || n1 < S::P) {  // only the last column can lie past P
#pragma unroll
            for (int n2 = 0; n2 < M; ++n2) c[i][n2] = src[(unsigned)(n1 + n2 * S::P)];
            w[i] = twn[(unsigned)n1];
        }
    }
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int n1 = j + i * S::kThreads;
        if (i + 1 < NI || n1 < S::P) {
            float2 a[M];
#pragma unroll
            for (int n2 = 0; n2 < M; ++n2) a[n2] = make_float2((float)c[i][n2].x, (float)c[i][n2].y);
            dft_odd<M>(a);
            lds[n1] = a[0];
            float2 wk = w[i];  // W_N^(n1 k2)
#pragma unroll
            for (int k2 = 1; k2 < M; ++k2) {
                lds[k2 * S::kSub + n1] = cmul(a[k2], wk);
                if (k2 + 1 < M) wk = cmul(wk, w[i]);
            }
        }
    }
}

// Z of one frame into the M spans (span k2 holds Z[M k1 + k2] at k1, plain
// layout); ends with a barrier. twp[m] = exp(-2 pi i m/P), m < P.
template <int M, int LOG2P, class Sync>
CORR_HD CORR_FI void mixed_fft_block(int j, const char2 *src, float2 *lds, Sync &sync, const float2 *twn,
                                    const float2 *twp) {
    using S = MixedSizes<M, LOG2P>;
    mixed_first_stage<M, LOG2P>(j, src, lds, twn);
    sync();
    const int s = j / S::kSubThreads, jl = j % S::kSubThreads;
    float2 *sub = lds + s * S::kSub;
    float2 v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = sub[jl + r * S::kSubThreads];
    sync();
    fft_from_values<LOG2P>(jl, v, sub, sync, twp);
}

// Z[k] of the mixed layout.
template <int M, int LOG2P>
CORR_HD CORR_FI float2 mixed_z(const float2 *lds, int k) {
    return lds[(unsigned)k % (unsigned)M * (unsigned)MixedSizes<M, LOG2P>::kSub + (unsigned)k / (unsigned)M];
}

// The R2C bins from a = Z[k], b = Z[N-k] (N-k taken mod N) and w = W^k:
// X[k] = E + w O, X[N-k] = conj(E - w O).
CORR_HD CORR_FI void r2c_from(float2 a, float2 b, float2 w, float2 &xk, float2 &xm) {
    const float er = 0.5f * (a.x + b.x), ei = 0.5f * (a.y - b.y);
    const float orr = 0.5f * (a.y + b.y), oi = -0.5f * (a.x - b.x);
    const float tr = w.x * orr - w.y * oi, ti = w.x * oi + w.y * orr;
    xk = make_float2(er + tr, ei + ti);
    xm = make_float2(er - tr, ti - ei);
}

// The R2C epilogue over the mixed layout, bins in the order and lane
// assignment of r2c_epilogue (k = j + q*N/16 with its mirror; bins 0, N/2).
template <int M, int LOG2P, class Emit>
CORR_HD CORR_FI void r2c_epilogue_mixed(int j, const float2 *lds, const float2 *wl, Emit &&emit) {
    constexpr int N = M << LOG2P;
    constexpr int NT = N / 16;
    const float2 wj = wl[(unsigned)j];
    emit.begin();
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int k = j + q * NT;
        float2 xk, xm;
        if (k == 0) {
            const float2 z0 = mixed_z<M, LOG2P>(lds, 0);
            r2c_from(z0, z0, wl[0], xk, xm);
            emit.at(0, xk);
        } else {
            r2c_combine(mixed_z<M, LOG2P>(lds, k), mixed_z<M, LOG2P>(lds, N - k), half_twiddle(wj, q), xk, xm);
            emit.pair(q, xk, xm);
        }
    }
    if (j == 0) {
        const float2 zh = mixed_z<M, LOG2P>(lds, N / 2);
        float2 xk, xm;
        r2c_from(zh, zh, wl[(unsigned)(N / 2)], xk, xm);
        emit.at(N / 2, xk);
    }
}

// One frame of one station, n_freq = M * 2^LOG2P; tw holds exp(-2 pi i m/N)
// for m < N followed by exp(-2 pi i m/P) for m < P.
template <int M, int LOG2P, class Sync, class Emit>
CORR_HD CORR_FI void fengine_mixed_frame(int j, int f, int st, float2 *lds, Sync &sync, const int8_t *in,
                                         size_t station_pitch, const int32_t *valid, size_t valid_stride,
                                         const float2 *tw, const float2 *wl, Emit &&emit) {
    constexpr int N = M << LOG2P;
    if (!frame_present(valid, valid_stride, f, st)) {
        emit_zero_frame<N>(j, emit);
        return;
    }
    const char2 *src =
        reinterpret_cast<const char2 *>(in + (size_t)st * station_pitch + (size_t)f * (size_t)(2 * N));
    mixed_fft_block<M, LOG2P>(j, src, lds, sync, tw, tw + N);
    r2c_epilogue_mixed<M, LOG2P>(j, lds, wl, emit);
}

template <int M, int LOG2P, int FORM, class Sync>
CORR_HD CORR_FI void fengine_mixed_form_block(int j, int f, int st, float2 *lds, Sync &sync, const int8_t *in,
                                             size_t station_pitch, const int32_t *valid, size_t valid_stride,
                                             const float2 *tw, const float2 *wl, const float *gain, void *out,
                                             size_t out_station_stride, int tile);

// Complex64 output, as fengine_fp32_block: tile = N is the plain order
// [station][frame][N], 16 or 32 the tile-major order of forms::spec_index.
template <int M, int LOG2P, class Sync>
CORR_HD CORR_FI void fengine_mixed_fp32_block(int j, int f, int st, float2 *lds, Sync &sync, const int8_t *in,
                                             size_t station_pitch, const int32_t *valid, size_t valid_stride,
                                             const float2 *tw, const float2 *wl, float2 *out,
                                             size_t out_station_stride, int tile = M << LOG2P) {
    fengine_mixed_form_block<M, LOG2P, 0>(j, f, st, lds, sync, in, station_pitch, valid, valid_stride, tw, wl,
                                          nullptr, out, out_station_stride, tile);
}

// The FP32 (FORM 0), FP16 and int8 forms, as fengine_form_block, with the
// spectrum tile given in channels: 16, 32 or N (the plain order). For N the emitter's tile shift
// is 15, beyond every bin, so that off(k) = k.
template <int M, int LOG2P, int FORM, class Sync>
CORR_HD CORR_FI void fengine_mixed_form_block(int j, int f, int st, float2 *lds, Sync &sync, const int8_t *in,
                                             size_t station_pitch, const int32_t *valid, size_t valid_stride,
                                             const float2 *tw, const float2 *wl, const float *gain, void *out,
                                             size_t out_station_stride, int tile) {
    constexpr int N = M << LOG2P;
    static_assert(N < (1 << 15), "the plain order's tile shift");
    using Emit = FormEmit<N, FORM>;
    const unsigned shift = tile == 16 ? 4u : tile == 32 ? 5u : 15u;
    Emit e;
    e.dst = static_cast<typename Emit::word_t *>(out) + (size_t)st * out_station_stride + (size_t)f * (size_t)tile;
    e.g = FORM == 0 ? gain : gain + (size_t)st * N;
    e.shift = shift;
    e.mask = (1u << shift) - 1u;
    e.tile_stride = tile == N ? 0u : (unsigned)(out_station_stride / N) * (unsigned)tile;  // fcap * T
    e.j = j;
    fengine_mixed_frame<M, LOG2P>(j, f, st, lds, sync, in, station_pitch, valid, valid_stride, tw, wl, e);
}

// ---------------------------------------------------------------------------
// n_freq above 16384 (every multiple of 2048 up to 2^20; the reference accepts
// every multiple of 2048, corr-cuda/src/corr_cuda.cpp:500-509): one frame's
// transform no longer fits a workgroup's LDS: N = m P with P the largest power
// of two dividing N, at most 16384, and m = N / P (2 to 511), and
//   Z[m k1 + k2] = DFT_P(y_k2)[k1],
//   y_k2[n1] = W_N^(n1 k2) sum_n2 z[n1 + P n2] W_m^(n2 k2)
// with W_m^r = W_N^(r P), the phase n2 k2 reduced mod m exactly (m complex
// products per point). Two forms, both from staged_sub_transform:
//  - one kernel (staged_frame): a workgroup transforms the sub-sequences of two
//    residues k2 that mirror each other under k -> N - k, one after the other
//    in LDS, and writes their R2C bins; nothing is staged, but the bins of one
//    workgroup lie m apart, so its stores scatter as m grows;
//  - two kernels: a workgroup per sub-sequence stores Z_k2 to a staging buffer
//    ([station][frame][m][P] complex64 of a batch of frames), and a lane per
//    bin pair reads Z[k] and Z[N-k] back and writes the bins in order.
// A missing frame's bins are written as zeros.
// ---------------------------------------------------------------------------
struct StagedPlan {
    int n = 0, m = 0, p = 0, log2p = 0;
};

// Largest n_freq of the staged transform: a frame of 2 MiB samples per input.
constexpr int kMaxStagedFreq = 1 << 20;

// The plan of an n_freq above 16384, or m = 0 where there is none.
CORR_HD inline StagedPlan staged_plan(int n_freq) {
    StagedPlan s;
    if (n_freq <= 16384 || n_freq > kMaxStagedFreq || n_freq % 2048 != 0) return s;
    int p = n_freq & -n_freq;
    if (p > 16384) p = 16384;
    s.n = n_freq;
    s.p = p;
    s.m = n_freq / p;
    while ((1 << s.log2p) < p) ++s.log2p;
    return s;
}

// The one-kernel form where it measured faster than the two-kernel one (2048
// frames of 24 inputs): the FP32 form at P = 16384 and m = 2 or 4 (-35% and
// -15%; +4% at m = 3, +11-19% at P <= 4096), the FP16 and int8 forms up to
// m = 11 (-13% at 18432 = 9 x 2048, -4% at 45056 = 11 x 4096, -35% at 32768,
// -25% at 65536; +33% at 67584 = 33 x 2048). Where it loses, its stores
// (a workgroup's bins lie m apart) cost more than the staging round trip.
CORR_HD inline bool staged_one_kernel(int form, const StagedPlan &pl) {
    if (form == 0) return pl.p == 16384 && (pl.m == 2 || pl.m == 4);
    return pl.m <= 11;
}

// Workgroup id -> (residue pair q, unit u = st * frames + f) on the one-kernel
// form's one-dimensional grid of X = m/2 + 1 workgroups for each of U units.
// The X workgroups of a unit get ids 8 apart: dispatched round robin over the
// 8 XCDs they share one XCD's L2, where their bins (interleaved m apart) merge
// into whole lines before being written back. The last U mod 8 units take the
// ids left, X in a row. A bijection of [0, X U) onto [0, X) x [0, U).
CORR_HD CORR_FI void staged_unit(unsigned id, unsigned X, unsigned U, unsigned &q, unsigned &u) {
    const unsigned full = U / 8u * 8u * X;
    if (id < full) {
        const unsigned b = id / (8u * X), r = id - b * 8u * X;
        q = r / 8u;
        u = b * 8u + r % 8u;
    } else {
        const unsigned r = id - full;
        q = r % X;
        u = U / 8u * 8u + r / X;
    }
}

// The P-point transform of sub-sequence k2 of the frame whose samples start at
// z, lane j of P/16: for its points n1 = j + r P/16,
//   y_k2[n1] = W_N^(n1 k2) sum_n2 z[n1 + P n2] W_m^(n2 k2),
// then the Stockham passes. On return lds holds Z_k2[k1] = Z[m k1 + k2] at k1
// (plain layout) after a final barrier.
template <int LOG2P, class Sync>
CORR_HD CORR_FI void staged_sub_transform(int j, int k2, const char2 *z, const StagedPlan &pl, float2 *lds,
                                         Sync &sync, const float2 *twn, const float2 *twp) {
    constexpr int P = 1 << LOG2P, NT = P / 16;
    float2 acc[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = make_float2(0.0f, 0.0f);
    int rr = 0;  // n2 * k2 mod m
    for (int n2 = 0; n2 < pl.m; ++n2) {
        const float2 w = twn[(unsigned)(rr * pl.p)];  // W_m^(n2 k2), the same for every lane
        char2 c[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) c[r] = z[(unsigned)(j + r * NT + n2 * P)];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            acc[r].x += (float)c[r].x * w.x - (float)c[r].y * w.y;
            acc[r].y += (float)c[r].x * w.y + (float)c[r].y * w.x;
        }
        rr += k2;
        if (rr >= pl.m) rr -= pl.m;
    }
    float2 v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = cmul(acc[r], twn[(unsigned)((j + r * NT) * k2)]);
    fft_from_values<LOG2P>(j, v, lds, sync, twp);
}

// Two-kernel form, stage 1: the workgroup of sub-sequence k2 of frame f of
// station st transforms it and stores Z_k2 to zs at ((st * frames + f) * m +
// k2) * P (the staging buffer of one batch of frames).
template <int LOG2P, class Sync>
CORR_HD CORR_FI void staged_first_sub(int j, int k2, int f, int st, int frames, const StagedPlan &pl,
                                     const int8_t *in, size_t station_pitch, int frame0, const int32_t *valid,
                                     size_t valid_stride, float2 *lds, Sync &sync, const float2 *twn,
                                     const float2 *twp, float2 *zs) {
    constexpr int P = 1 << LOG2P, NT = P / 16;
    if (!frame_present(valid, valid_stride, frame0 + f, st)) return;  // uniform over the workgroup
    const char2 *z = reinterpret_cast<const char2 *>(in + (size_t)st * station_pitch +
                                                     (size_t)(frame0 + f) * (size_t)(2 * pl.n));
    staged_sub_transform<LOG2P>(j, k2, z, pl, lds, sync, twn, twp);
    const size_t base = (((size_t)st * frames + f) * pl.m + k2) * P;
#pragma unroll
    for (int r = 0; r < 16; ++r) zs[base + (size_t)(j + r * NT)] = lds[j + r * NT];
}

// Two-kernel form, stage 2: lane k (0 <= k <= N/2) of frame f of station st: bins k and N-k
// (k alone for 0 and N/2) from Z[k] = zs[k mod m][k div m]. `put(k, X)`
// stores bin k, `zero(k)` a missing frame's bin.
template <class Put, class Zero>
CORR_HD CORR_FI void staged_r2c(int k, int f, int st, int frames, const StagedPlan &pl, int frame0,
                               const int32_t *valid, size_t valid_stride, const float2 *zs, const float2 *wl,
                               Put &&put, Zero &&zero) {
    const int n = pl.n;
    if (k > n / 2) return;
    const bool pair = k != 0 && k != n / 2;
    if (!frame_present(valid, valid_stride, frame0 + f, st)) {
        zero(k);
        if (pair) zero(n - k);
        return;
    }
    const float2 *zf = zs + ((size_t)st * frames + f) * (size_t)n;
    const int km = n - k == n ? 0 : n - k;
    const float2 a = zf[(size_t)(k % pl.m) * pl.p + (size_t)(k / pl.m)];
    const float2 b = zf[(size_t)(km % pl.m) * pl.p + (size_t)(km / pl.m)];
    float2 xk, xm;
    r2c_from(a, b, wl[(unsigned)k], xk, xm);
    put(k, xk);
    if (pair) put(n - k, xm);
}

// One-kernel form: workgroup q (0 <= q <= m/2) of frame f of station st, lane j of P/16: every
// bin k = m k1 + c of the residues c = a = q and c = b = (m - q) mod m, written
// by put(k, X) (zero(k) for a missing frame). Bin k needs Z[k] and Z[N-k], and
// N - (m k1 + a) = m (P-1-k1) + b for a > 0, so the workgroup transforms
// sub-sequence a, keeps its spectrum in registers (16 points a lane), transforms
// b in the same LDS and forms each pair (k, N-k) from the two. For a = b (the
// residue 0, and m/2 for even m) one transform serves, each mirror pair formed
// once; bins 0 and N/2 (residue 0, k1 = 0 and P/2) are alone.
// Each of the frame's m/2 + 1 workgroups reads all of the frame's samples (2N
// bytes, m times the input in all, from L2 for the most part); nothing but the
// spectra is written.
template <int LOG2P, class Sync, class Put, class Zero>
CORR_HD CORR_FI void staged_frame(int j, int q, int f, int st, const StagedPlan &pl, const int8_t *in,
                                 size_t station_pitch, int frame0, const int32_t *valid, size_t valid_stride,
                                 float2 *lds, Sync &sync, const float2 *twn, const float2 *twp, const float2 *wl,
                                 Put &&put, Zero &&zero) {
    constexpr int P = 1 << LOG2P, NT = P / 16;
    const int m = pl.m, n = pl.n, a = q, b = q == 0 ? 0 : m - q;
    if (!frame_present(valid, valid_stride, frame0 + f, st)) {  // uniform over the workgroup
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int k1 = j + r * NT;
            zero(m * k1 + a);
            if (b != a) zero(m * k1 + b);
        }
        return;
    }
    const char2 *z = reinterpret_cast<const char2 *>(in + (size_t)st * station_pitch +
                                                     (size_t)(frame0 + f) * (size_t)(2 * n));
    staged_sub_transform<LOG2P>(j, a, z, pl, lds, sync, twn, twp);
    if (b != a) {
        float2 za[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) za[r] = lds[j + r * NT];
        sync();  // every lane holds its points of Z_a before Z_b overwrites them
        staged_sub_transform<LOG2P>(j, b, z, pl, lds, sync, twn, twp);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int k1 = j + r * NT, k = m * k1 + a;
            float2 xk, xm;
            r2c_from(za[r], lds[P - 1 - k1], wl[(unsigned)k], xk, xm);
            put(k, xk);
            put(n - k, xm);
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int k1 = j + r * NT;
        if (a == 0 ? k1 > P / 2 : k1 >= P / 2) continue;
        const int k1m = a == 0 ? (P - k1) & (P - 1) : P - 1 - k1, k = m * k1 + a;
        float2 xk, xm;
        r2c_from(lds[k1], lds[k1m], wl[(unsigned)k], xk, xm);
        put(k, xk);
        if (k1m != k1) put(n - k, xm);
    }
}

// Calibration: power[st*nf + c] += sum over the frames present of |Y|^2, in
// frame order; count[st] += frames present (by the channel-0 thread).
// Thread c of the workgroup row of station st.
CORR_HD CORR_FI void power_sum(int c, int st, int nf, int frames, const float2 *spec, size_t spec_station_stride,
                              const int32_t *valid, size_t valid_stride, float *power, int32_t *count) {
    if (c >= nf) return;
    const float2 *y = spec + (size_t)st * spec_station_stride + (size_t)c;
    float p = 0.0f;
    int n = 0;
    for (int t = 0; t < frames; ++t) {
        if (valid != nullptr && valid[(size_t)st * valid_stride + (size_t)t] == 0) continue;
        const float2 v = y[(size_t)t * (size_t)nf];
        p += v.x * v.x + v.y * v.y;
        ++n;
    }
    power[(size_t)st * nf + c] += p;
    if (c == 0) count[st] += n;
}

// Gain of (station st, channel c) from the calibration power (forms.hpp):
//   FP16: 2^(8-E) with sigma = m 2^E, m in [0.5, 1): the component RMS lands in [128, 256)
//   int8: 1/(0.35 sigma)
// sigma = per-component RMS = sqrt(mean |Y|^2 / 2); gain 1 where nothing was seen.
CORR_HD CORR_FI float gain_from_power(int form, float power, int32_t count) {
    if (count <= 0 || !(power > 0.0f) || !(power < 3.0e38f)) return 1.0f;
    const float sigma = sqrtf(power / (float)count * 0.5f);
    if (!(sigma > 0.0f)) return 1.0f;
    if (form == 1) {
        int e = 0;
        (void)frexpf(sigma, &e);
        int s = 8 - e;
        if (s > 60) s = 60;
        if (s < -60) s = -60;
        return ldexpf(1.0f, s);
    }
    return 1.0f / (forms::kInt8StepSigma * sigma);
}

}  // namespace fft
}  // namespace corr
