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

namespace corr {
namespace {

// ---------------------------------------------------------------------------
// small complex helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 w) {
    return make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}

// cos(2 pi m / 32), m = 0..8; the rest by symmetry.
__device__ constexpr float kCosOct[9] = {
    1.0f,
    0.98078528040323044913f,
    0.92387953251128675613f,
    0.83146961230254523708f,
    0.70710678118654752440f,
    0.55557023301960222474f,
    0.38268343236508977173f,
    0.19509032201612826785f,
    0.0f,
};

__device__ __forceinline__ constexpr float cos32(int m) {
    return m <= 8 ? kCosOct[m] : m <= 16 ? -kCosOct[16 - m] : m <= 24 ? -kCosOct[m - 16] : kCosOct[32 - m];
}

// x * exp(-2 pi i m / 32) for a compile-time m in [0, 16).
template <int M>
__device__ __forceinline__ float2 rot32(float2 x) {
    if constexpr (M == 0) {
        return x;
    } else if constexpr (M == 8) {  // -i
        return make_float2(x.y, -x.x);
    } else if constexpr (M == 4) {  // (1 - i)/sqrt2
        constexpr float r = kCosOct[4];
        return make_float2((x.x + x.y) * r, (x.y - x.x) * r);
    } else if constexpr (M == 12) {  // (-1 - i)/sqrt2
        constexpr float r = kCosOct[4];
        return make_float2((x.y - x.x) * r, -(x.x + x.y) * r);
    } else {
        constexpr float c = cos32(M);
        constexpr float s = cos32((M + 24) & 31);  // sin(2 pi M / 32)
        return make_float2(x.x * c + x.y * s, x.y * c - x.x * s);
    }
}

// One radix-2 DIF stage of an in-register DFT of size R: blocks of 2*H.
template <int R, int H, int BASE, int K>
__device__ __forceinline__ void dif_butterflies(float2 *v) {
    if constexpr (BASE < R) {
        if constexpr (K < H) {
            const float2 a = v[BASE + K], b = v[BASE + K + H];
            v[BASE + K] = cadd(a, b);
            // W_{2H}^K = exp(-2 pi i K/(2H)) = rot32 by K*32/(2H)
            v[BASE + K + H] = rot32<K * (32 / (2 * H))>(csub(a, b));
            dif_butterflies<R, H, BASE, K + 1>(v);
        } else {
            dif_butterflies<R, H, BASE + 2 * H, 0>(v);
        }
    }
}

template <int R, int H>
__device__ __forceinline__ void dif_stages(float2 *v) {
    if constexpr (H >= 1) {
        dif_butterflies<R, H, 0, 0>(v);
        dif_stages<R, H / 2>(v);
    }
}

template <int R>
__host__ __device__ constexpr int bitrev(int i) {
    int r = 0;
    for (int b = 1; b < R; b <<= 1) {
        r = (r << 1) | (i & 1);
        i >>= 1;
    }
    return r;
}

template <int R, int I>
__device__ __forceinline__ void unscramble(float2 *v) {
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

// In-place DFT of R (power of two, <= 32) values, natural order in and out:
// v[k] := sum_r v[r] exp(-2 pi i r k / R).
template <int R>
__device__ __forceinline__ void dft(float2 *v) {
    dif_stages<R, R / 2>(v);
    unscramble<R, 0>(v);
}

// LDS index padding: one spare float2 every 16, so that a wave writing with a
// stride of 16 elements (the first Stockham pass) spreads over all banks.
__device__ __forceinline__ int pad16(int a) { return a + (a >> 4); }

// One Stockham pass of radix R over N points, Ns = points already combined,
// reading and writing the LDS buffer in place (barrier between all reads and
// any write). Thread j does 16/R butterflies jj = j + q*N/16.
template <int N, int R, int NS>
__device__ __forceinline__ void stockham_pass(float2 (&v)[16], float2 *lds,
                                              const float2 *__restrict__ tw, int j) {
    constexpr int NT = N / 16;
    constexpr int NB = 16 / R;
    constexpr int STRIDE = N / R;
    constexpr int TWSTEP = N / (NS * R);
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int jj = j + q * NT;
#pragma unroll
        for (int r = 0; r < R; ++r) v[q * R + r] = lds[pad16(jj + r * STRIDE)];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int jj = j + q * NT;
        const int k = jj & (NS - 1);
#pragma unroll
        for (int r = 1; r < R; ++r) v[q * R + r] = cmul(v[q * R + r], tw[r * k * TWSTEP]);
        dft<R>(&v[q * R]);
        const int base = (jj - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[pad16(base + r * NS)] = v[q * R + r];
    }
    __syncthreads();
}

template <int LOG2N, int DONE>
__device__ __forceinline__ void stockham_rest(float2 (&v)[16], float2 *lds,
                                              const float2 *__restrict__ tw, int j) {
    if constexpr (DONE < LOG2N) {
        constexpr int REM = LOG2N - DONE;
        constexpr int LOGR = REM >= 4 ? 4 : REM;
        stockham_pass<(1 << LOG2N), (1 << LOGR), (1 << DONE)>(v, lds, tw, j);
        stockham_rest<LOG2N, DONE + LOGR>(v, lds, tw, j);
    }
}

// ---------------------------------------------------------------------------
// F-engine, complex64 output
// grid (frames, stations), block N/16 threads, N = n_freq = 2^LOG2N
// ---------------------------------------------------------------------------
template <int LOG2N>
__global__ void __launch_bounds__((1 << LOG2N) / 16)
fengine_fp32_kernel(const int8_t *__restrict__ in, size_t station_pitch,
                    const uint8_t *__restrict__ mask, size_t mask_pitch, int mask_bit0,
                    const float2 *__restrict__ tw, const float2 *__restrict__ wl,
                    float2 *__restrict__ out, size_t out_station_stride) {
    constexpr int N = 1 << LOG2N;
    constexpr int NT = N / 16;
    __shared__ float2 lds[N + N / 16];

    const int j = threadIdx.x;
    const int f = blockIdx.x;
    const int st = blockIdx.y;

    bool present = true;
    if (mask != nullptr) {
        const int bit = mask_bit0 + f;
        present = (mask[(size_t)st * mask_pitch + (size_t)(bit >> 3)] >> (bit & 7)) & 1u;
    }

    float2 v[16];
    if (present) {
        const char2 *src = reinterpret_cast<const char2 *>(in + (size_t)st * station_pitch +
                                                           (size_t)f * (size_t)(2 * N));
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const char2 p = src[j + r * NT];
            v[r] = make_float2((float)p.x, (float)p.y);
        }
    } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = make_float2(0.0f, 0.0f);
    }

    // first pass: radix 16, Ns = 1, inputs z[j + r*N/16] already in registers
    dft<16>(v);
#pragma unroll
    for (int r = 0; r < 16; ++r) lds[pad16(j * 16 + r)] = v[r];
    __syncthreads();
    stockham_rest<LOG2N, 4>(v, lds, tw, j);

    // R2C unpack of the retained bins
    float2 *dst = out + (size_t)st * out_station_stride + (size_t)f * (size_t)N;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int k = j + q * NT;
        const float2 a = lds[pad16(k)];
        const float2 b = lds[pad16((N - k) & (N - 1))];
        const float er = 0.5f * (a.x + b.x), ei = 0.5f * (a.y - b.y);
        const float orr = 0.5f * (a.y + b.y), oi = -0.5f * (a.x - b.x);
        const float2 w = wl[k];
        dst[k] = make_float2(er + (w.x * orr - w.y * oi), ei + (w.x * oi + w.y * orr));
    }
}

// Twiddle tables in FP64: tw[m] = exp(-2 pi i m/n), wl[k] = exp(-2 pi i k/(2n)).
__global__ void twiddle_kernel(float2 *tw, float2 *wl, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
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
                    int frames_lds, float2 *__restrict__ acc) {
    extern __shared__ float2 sy[];
    const int lane = (int)(threadIdx.x & 63);
    const int w = (int)(threadIdx.x >> 6);
    const int W = (int)(blockDim.x >> 6);
    const int c = (int)blockIdx.x * 64 + lane;
    const int npad = (n_station + 3) & ~3;

    int2 bp[NPAIR];
    bool act[NPAIR];
#pragma unroll
    for (int q = 0; q < NPAIR; ++q) {
        const int p = ((int)blockIdx.y * W + w) * NPAIR + q;
        act[q] = p < n_bpairs;
        bp[q] = act[q] ? bpairs[p] : make_int2(0, 0);
    }

    float2 a[NPAIR][16];
#pragma unroll
    for (int q = 0; q < NPAIR; ++q)
#pragma unroll
        for (int e = 0; e < 16; ++e) a[q][e] = make_float2(0.0f, 0.0f);

    for (int f0 = 0; f0 < frames; f0 += frames_lds) {
        const int ft = min(frames_lds, frames - f0);
        __syncthreads();
        for (int row = w; row < npad * ft; row += W) {
            const int i = row / ft, t = row - (row / ft) * ft;
            float2 val = make_float2(0.0f, 0.0f);
            if (i < n_station && c < nf) val = y[(size_t)i * ys + (size_t)(f0 + t) * nf + c];
            sy[(i * frames_lds + t) * 64 + lane] = val;
        }
        __syncthreads();
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
                        a[q][e * 4 + g].x += xj[e].x * xk[g].x + xj[e].y * xk[g].y;
                        a[q][e * 4 + g].y += xj[e].y * xk[g].x - xj[e].x * xk[g].y;
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

__global__ void scale_kernel(float2 *acc, size_t n, float s) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float2 v = acc[i];
        v.x *= s;
        v.y *= s;
        acc[i] = v;
    }
}

template <int LOG2N>
hipError_t launch_fengine_fp32_t(const int8_t *in, size_t station_pitch, const uint8_t *mask,
                                 size_t mask_pitch, int mask_bit0, const float2 *tw,
                                 const float2 *wl, float2 *out, size_t out_station_stride,
                                 int frames, int n_station, hipStream_t s) {
    const dim3 grid((unsigned)frames, (unsigned)n_station);
    const dim3 block((1u << LOG2N) / 16u);
    hipLaunchKernelGGL(fengine_fp32_kernel<LOG2N>, grid, block, 0, s, in, station_pitch, mask,
                       mask_pitch, mask_bit0, tw, wl, out, out_station_stride);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_twiddles(float2 *tw, float2 *wl, int n_freq, hipStream_t s) {
    hipLaunchKernelGGL(twiddle_kernel, dim3((unsigned)((n_freq + 255) / 256)), dim3(256), 0, s,
                       tw, wl, n_freq);
    return hipGetLastError();
}

hipError_t launch_fengine_fp32(int log2_nf, const int8_t *in, size_t station_pitch,
                               const uint8_t *mask, size_t mask_pitch, int mask_bit0,
                               const float2 *tw, const float2 *wl, float2 *out,
                               size_t out_station_stride, int frames, int n_station,
                               hipStream_t s) {
    if (frames <= 0 || n_station <= 0) return hipSuccess;
    switch (log2_nf) {
#define CORR_FENGINE_CASE(L)                                                                \
    case L:                                                                                 \
        return launch_fengine_fp32_t<L>(in, station_pitch, mask, mask_pitch, mask_bit0, tw, \
                                        wl, out, out_station_stride, frames, n_station, s);
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
    const int npad = (n_station + 3) & ~3;
    const size_t shm = (size_t)npad * (size_t)frames_lds * 64 * sizeof(float2);
    const dim3 grid((unsigned)((n_freq + 63) / 64), (unsigned)groups);
    const dim3 block((unsigned)(64 * waves));
    switch (npair) {
#define CORR_XENGINE_CASE(P)                                                                  \
    case P:                                                                                   \
        hipLaunchKernelGGL(xengine_fp32_kernel<P>, grid, block, shm, s, y, y_station_stride, \
                           n_freq, frames, n_station, bpairs, n_bpairs, frames_lds, acc);     \
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
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(scale_kernel, dim3((unsigned)blocks), dim3(256), 0, st, acc, n, s);
    return hipGetLastError();
}

}  // namespace corr
