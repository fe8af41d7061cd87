/*
 * corr_oracle.c -- CPU restatement of the reference correlator path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the shipped library links, loads or calls
 * this file: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it, and only as the checker (or, for the CPU baseline, as the thing timed
 * beside the GPU). The GPU path never falls back to it.
 *
 * What it restates (file:line into the reference):
 *   orc_pairs            baseline b <-> (j,k), j <= k, j outer
 *                        (corr-cuda/src/corr_cuda.cpp:193-204, 1515-1534)
 *   orc_gen_int8         the seeded correlated-Gaussian int8 generator of the
 *                        benchmark configurations (the reference has none; its own
 *                        test input is numpy's, corr-cuda/tests/gen_golden.py:59-62)
 *   orc_fft_r2c_f64      unnormalised R2C of length L = 2*n_freq, Nyquist dropped
 *                        (cuFFT plan corr-cuda/src/kernels.cu:578-592, exec :832-835;
 *                        numpy restatement corr-cuda/tests/gen_golden.py:68-69)
 *   orc_dft_r2c_f64      the same by brute-force DFT (tiny L only)
 *   orc_fx_f64           mask-zero + exact int8->double + R2C + sum_t Y_j conj(Y_k)
 *                        + /frames, all in double with compensated sums
 *                        (kernels.cu:130-139, 159-174, 243-246, 178-185;
 *                        gen_golden.py:71-78)
 *   orc_fx_f32           the reference's FP32 arithmetic and operation order:
 *                        float conversion, FP32 FFT, per-chunk frame-order FP32
 *                        partial sums added into the row once per chunk, then
 *                        row *= 1.0f/(float)frames (kernels.cu:216-264, 890-891).
 *                        This is the CPU path timed as the baseline.
 *   orc_x_f64_from_spectra  X-engine in double on given (e.g. requantized) spectra
 *
 * Pinning (how each function is checked; see tests/test_oracle.py):
 *   orc_pairs         abi_smoke.c:86-93 known answer (4 stations), tests/abi.rs
 *                     counts for N=1..24, closed-form index j*N - j(j-1)/2 + (k-j)
 *   orc_fft_r2c_f64   brute-force DFT (tiny L, powers of two and odd parts 3-15),
 *                     numpy.fft.rfft at L=2048..32768 and at the non-power-of-two
 *                     L = 2 * {6144, 10240, 12288, 14336, 18432},
 *                     impulse -> flat, pure tone -> single bin, Parseval
 *   orc_fx_f64        numpy FP64 restatement of gen_golden.py on committed
 *                     fixtures; autocorrelation imaginary part ~ 0; delay test
 *                     phase slope 2*pi*tau*c/L (tools/phase_check.py:4-15)
 *   orc_fx_f32        relative RMS vs orc_fx_f64 below the reference's golden
 *                     tolerance 1e-4 (corr-cuda/tests/golden.c:41-48)
 *   orc_gen_int8      SHA-256 of its output committed per configuration, plus
 *                     moment checks (sigma, rho, delay recovered by the
 *                     correlator's phase slope)
 */

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846264338327950288

/* ------------------------------------------------------------------------- */
/* Baseline numbering                                                         */
/* ------------------------------------------------------------------------- */

int orc_n_baseline(int n) { return n < 1 ? -1 : n * (n + 1) / 2; }

void orc_pairs(int n, int32_t *out_jk) {
    int i = 0;
    for (int j = 0; j < n; ++j)
        for (int k = j; k < n; ++k) {
            out_jk[2 * i] = j;
            out_jk[2 * i + 1] = k;
            ++i;
        }
}

/* ------------------------------------------------------------------------- */
/* Seeded generator                                                           */
/*                                                                            */
/* x_i[t] = clip(round(sigma*(sqrt(rho)*s[t-d_i] + sqrt(1-rho)*n_i[t])        */
/*                     + A*cos(2*pi*c*t/L)), -127, 127)                       */
/* s, n_i i.i.d. N(0,1) from a counter-based hash (splitmix64 finaliser) and  */
/* Box-Muller; d_i in [0, max_delay] from the same hash. round() is C's       */
/* round-half-away-from-zero; t is the global sample index frame*L + n.       */
/* ------------------------------------------------------------------------- */

static uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint64_t ctr_hash(uint64_t seed, uint64_t stream, uint64_t idx) {
    const uint64_t key = mix64(seed * 0x9E3779B97F4A7C15ull + stream * 0xD1B54A32D192ED03ull);
    return mix64(key ^ mix64(idx + 0x632BE59BD9B4E019ull));
}

static double u01(uint64_t v) { return ((double)(v >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

/* Standard normal number `idx` of stream `stream`. */
static double gauss(uint64_t seed, uint64_t stream, uint64_t idx) {
    const double u1 = u01(ctr_hash(seed, stream, 2 * idx));
    const double u2 = u01(ctr_hash(seed, stream, 2 * idx + 1));
    return sqrt(-2.0 * log(u1)) * cos(2.0 * ORC_PI * u2);
}

#define ORC_STREAM_COMMON 0ull
#define ORC_STREAM_DELAY 100000ull
#define ORC_STREAM_NOISE0 1ull

void orc_delays(uint64_t seed, int n_station, int max_delay, int32_t *out) {
    for (int i = 0; i < n_station; ++i)
        out[i] = (int32_t)(ctr_hash(seed, ORC_STREAM_DELAY, (uint64_t)i) %
                           (uint64_t)(max_delay + 1));
}

/* out: int8 [n_station][n_frames][L], frames starting at global frame `frame0`. */
void orc_gen_int8(uint64_t seed, int n_station, int64_t frame0, int n_frames, int L,
                  double rho, double sigma, int max_delay, int tone_channel,
                  double tone_amp, int8_t *out) {
    int32_t *d = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n_station > 0 ? n_station : 1));
    orc_delays(seed, n_station, max_delay, d);
    const double a = sqrt(rho), b = sqrt(1.0 - rho);
    const int64_t total = (int64_t)n_frames * L;
#pragma omp parallel for schedule(static) collapse(2)
    for (int i = 0; i < n_station; ++i) {
        for (int64_t u = 0; u < total; ++u) {
            const int64_t t = frame0 * (int64_t)L + u;
            /* s is indexed from -max_delay so that t - d_i is never negative. */
            const double s = gauss(seed, ORC_STREAM_COMMON, (uint64_t)(t - d[i] + max_delay));
            const double nz = gauss(seed, ORC_STREAM_NOISE0 + (uint64_t)i, (uint64_t)t);
            double v = sigma * (a * s + b * nz);
            if (tone_amp != 0.0) {
                const int64_t ph = (t % L) * (int64_t)tone_channel % L;
                v += tone_amp * cos(2.0 * ORC_PI * (double)ph / (double)L);
            }
            double r = round(v);
            if (r > 127.0) r = 127.0;
            if (r < -127.0) r = -127.0;
            out[(size_t)i * (size_t)total + (size_t)u] = (int8_t)r;
        }
    }
    free(d);
}

/* ------------------------------------------------------------------------- */
/* FP64 FFT: iterative radix-2 on the full L-point complex sequence (x, 0).   */
/* Deliberately a different algorithm from the GPU's half-length packing.     */
/* ------------------------------------------------------------------------- */

static void fft_c64(double *re, double *im, int n) {
    /* bit reversal */
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1;
        for (int k = 0; k < half; ++k) {
            /* twiddle computed directly, not by recurrence */
            const double ang = -2.0 * ORC_PI * (double)k / (double)len;
            const double wr = cos(ang), wi = sin(ang);
            for (int s = 0; s < n; s += len) {
                const int p = s + k, q = p + half;
                const double xr = re[q] * wr - im[q] * wi;
                const double xi = re[q] * wi + im[q] * wr;
                re[q] = re[p] - xr; im[q] = im[p] - xi;
                re[p] += xr;        im[p] += xi;
            }
        }
    }
}

/* Any length n = m * 2^a: decimation in time by the odd part m. The m
   sub-sequences x[m t + r] (t < 2^a) are transformed by fft_c64, then
   X[k] = sum_r W_n^(r k) X_r[k mod 2^a], a direct m-term sum per bin with the
   phase reduced exactly, (r k) mod n, into a table of W_n built in double.
   The reference accepts every n_freq that is a multiple of 2048
   (corr-cuda/src/corr_cuda.cpp:500-509), so L = 2 n_freq has an odd part of
   any size; the work is O(n (m + log n)). */
static void fft_any_c64(double *re, double *im, int n) {
    const int p = n & -n;
    const int m = n / p;
    if (m == 1) {
        fft_c64(re, im, n);
        return;
    }
    double *sr = (double *)malloc(sizeof(double) * (size_t)n);
    double *si = (double *)malloc(sizeof(double) * (size_t)n);
    double *wr = (double *)malloc(sizeof(double) * (size_t)n);
    double *wi = (double *)malloc(sizeof(double) * (size_t)n);
    for (int r = 0; r < m; ++r)
        for (int t = 0; t < p; ++t) {
            sr[(size_t)r * p + t] = re[(size_t)m * t + r];
            si[(size_t)r * p + t] = im[(size_t)m * t + r];
        }
    if (p > 1)
        for (int r = 0; r < m; ++r) fft_c64(sr + (size_t)r * p, si + (size_t)r * p, p);
    for (int q = 0; q < n; ++q) {
        const double ang = -2.0 * ORC_PI * (double)q / (double)n;
        wr[q] = cos(ang);
        wi[q] = sin(ang);
    }
    for (int k = 0; k < n; ++k) {
        double ar = 0.0, ai = 0.0;
        for (int r = 0; r < m; ++r) {
            const int q = (int)(((long long)r * k) % n);
            const double xr = sr[(size_t)r * p + (k & (p - 1))], xi = si[(size_t)r * p + (k & (p - 1))];
            ar += xr * wr[q] - xi * wi[q];
            ai += xr * wi[q] + xi * wr[q];
        }
        re[k] = ar;
        im[k] = ai;
    }
    free(sr);
    free(si);
    free(wr);
    free(wi);
}

/* x: L real doubles; out: (L/2) complex as interleaved (re, im); bins 0..L/2-1. */
void orc_fft_r2c_f64(const double *x, int L, double *out) {
    double *re = (double *)malloc(sizeof(double) * (size_t)L);
    double *im = (double *)malloc(sizeof(double) * (size_t)L);
    for (int i = 0; i < L; ++i) { re[i] = x[i]; im[i] = 0.0; }
    fft_any_c64(re, im, L);
    for (int k = 0; k < L / 2; ++k) { out[2 * k] = re[k]; out[2 * k + 1] = im[k]; }
    free(re);
    free(im);
}

/* Brute-force DFT, O(L^2); phase reduced exactly as (k*t) mod L. */
void orc_dft_r2c_f64(const double *x, int L, double *out) {
    for (int k = 0; k < L / 2; ++k) {
        long double sr = 0.0L, si = 0.0L;
        for (int t = 0; t < L; ++t) {
            const long long ph = ((long long)k * t) % L;
            const long double ang = -2.0L * (long double)ORC_PI * (long double)ph / (long double)L;
            sr += (long double)x[t] * cosl(ang);
            si += (long double)x[t] * sinl(ang);
        }
        out[2 * k] = (double)sr;
        out[2 * k + 1] = (double)si;
    }
}

/* ------------------------------------------------------------------------- */
/* FP64 FX                                                                    */
/* ------------------------------------------------------------------------- */

/* Validity of frame f of station i: valid == NULL means every frame is present,
   else valid[i*n_frames + f] != 0. A missing frame is correlated as zeros
   (kernels.cu:159-174; corr_cuda.h:268-291). */
static int frame_ok(const uint8_t *valid, int i, int f, int n_frames) {
    return valid == NULL || valid[(size_t)i * (size_t)n_frames + (size_t)f] != 0;
}

/* x: int8 [N][T][L]; V out: complex double interleaved [B][nf], divided by
   `frames_div` (the caller's nominal n_cum, not the valid count). */
void orc_fx_f64(const int8_t *x, const uint8_t *valid, int N, int T, int L,
                double frames_div, double *V) {
    const int nf = L / 2;
    const int B = orc_n_baseline(N);
    double *Y = (double *)malloc(sizeof(double) * 2 * (size_t)N * (size_t)T * (size_t)nf);
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * (size_t)L);
#pragma omp for schedule(dynamic) collapse(2)
        for (int i = 0; i < N; ++i)
            for (int t = 0; t < T; ++t) {
                const int8_t *src = x + ((size_t)i * T + t) * (size_t)L;
                const int ok = frame_ok(valid, i, t, T);
                for (int n = 0; n < L; ++n) buf[n] = ok ? (double)src[n] : 0.0;
                orc_fft_r2c_f64(buf, L, Y + ((size_t)i * T + t) * 2 * (size_t)nf);
            }
        free(buf);
    }
    int32_t *pj = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)B);
    orc_pairs(N, pj);
#pragma omp parallel for schedule(dynamic) collapse(2)
    for (int b = 0; b < B; ++b)
        for (int c = 0; c < nf; ++c) {
            const int j = pj[2 * b], k = pj[2 * b + 1];
            long double sr = 0.0L, si = 0.0L;
            for (int t = 0; t < T; ++t) {
                const double *a = Y + (((size_t)j * T + t) * nf + c) * 2;
                const double *q = Y + (((size_t)k * T + t) * nf + c) * 2;
                sr += (long double)a[0] * q[0] + (long double)a[1] * q[1];
                si += (long double)a[1] * q[0] - (long double)a[0] * q[1];
            }
            V[((size_t)b * nf + c) * 2] = (double)(sr / frames_div);
            V[((size_t)b * nf + c) * 2 + 1] = (double)(si / frames_div);
        }
    free(pj);
    free(Y);
}

/* X-engine alone in double, on spectra given as float complex [N][T][nf]
   (e.g. the device's requantized spectra, expanded to float), with optional
   per-(input, channel) gains g[i*nf + c] that were applied before quantizing:
   V_jk = sum_t y_j conj(y_k) / (g_j g_k) / frames_div. */
void orc_x_f64_from_spectra(const float *Yf, const float *gain, int N, int T, int nf,
                            double frames_div, double *V) {
    const int B = orc_n_baseline(N);
    int32_t *pj = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)B);
    orc_pairs(N, pj);
#pragma omp parallel for schedule(dynamic) collapse(2)
    for (int b = 0; b < B; ++b)
        for (int c = 0; c < nf; ++c) {
            const int j = pj[2 * b], k = pj[2 * b + 1];
            long double sr = 0.0L, si = 0.0L;
            for (int t = 0; t < T; ++t) {
                const float *a = Yf + (((size_t)j * T + t) * nf + c) * 2;
                const float *q = Yf + (((size_t)k * T + t) * nf + c) * 2;
                sr += (long double)a[0] * q[0] + (long double)a[1] * q[1];
                si += (long double)a[1] * q[0] - (long double)a[0] * q[1];
            }
            long double g = 1.0L;
            if (gain != NULL)
                g = (long double)gain[(size_t)j * nf + c] * (long double)gain[(size_t)k * nf + c];
            V[((size_t)b * nf + c) * 2] = (double)(sr / g / frames_div);
            V[((size_t)b * nf + c) * 2 + 1] = (double)(si / g / frames_div);
        }
    free(pj);
}

/* ------------------------------------------------------------------------- */
/* FP32 path in the reference's arithmetic (the CPU baseline)                 */
/* ------------------------------------------------------------------------- */

typedef struct { float re, im; } cf32;

/* nf-point complex FFT in float, iterative radix-2, twiddles from a table
   computed in double. `tw` holds exp(-2 pi i m / nf), m < nf/2. */
static void fft_c32(cf32 *a, int n, const cf32 *tw) {
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cf32 t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int s = 0; s < n; s += len) {
            for (int k = 0; k < half; ++k) {
                const cf32 w = tw[k * step];
                cf32 *p = a + s + k, *q = p + half;
                const float xr = q->re * w.re - q->im * w.im;
                const float xi = q->re * w.im + q->im * w.re;
                q->re = p->re - xr; q->im = p->im - xi;
                p->re += xr;        p->im += xi;
            }
        }
    }
}

/* Unnormalised float R2C of L=2*nf real samples through an nf-point complex FFT
   of z[m] = x[2m] + i x[2m+1]; bins 0..nf-1. */
static void r2c_f32(const float *x, int nf, const cf32 *tw, const cf32 *wl, cf32 *z,
                    cf32 *out) {
    for (int m = 0; m < nf; ++m) { z[m].re = x[2 * m]; z[m].im = x[2 * m + 1]; }
    fft_c32(z, nf, tw);
    for (int k = 0; k < nf; ++k) {
        const cf32 a = z[k], b = z[(nf - k) & (nf - 1)];
        /* E = (Z[k] + conj(Z[nf-k]))/2, O = (Z[k] - conj(Z[nf-k]))/(2i) */
        const float er = 0.5f * (a.re + b.re), ei = 0.5f * (a.im - b.im);
        const float or_ = 0.5f * (a.im + b.im), oi = -0.5f * (a.re - b.re);
        out[k].re = er + (wl[k].re * or_ - wl[k].im * oi);
        out[k].im = ei + (wl[k].re * oi + wl[k].im * or_);
    }
}

/* x: int8 [N][T][L]; V out: complex float interleaved [B][nf].
   chunk_frames: frames per partial sum (kernels.cu:216-264 flushes once per
   chunk); frames_div: the caller's n_cum (kernels.cu:890-891). Returns the
   number of OpenMP threads used. */
int orc_fx_f32(const int8_t *x, const uint8_t *valid, int N, int T, int L,
               int chunk_frames, uint64_t frames_div, float *V) {
    const int nf = L / 2;
    const int B = orc_n_baseline(N);
    if (nf < 1 || (nf & (nf - 1)) != 0) return -1; /* its radix-2 FFT: power-of-two n_freq */
    cf32 *tw = (cf32 *)malloc(sizeof(cf32) * (size_t)(nf / 2 > 0 ? nf / 2 : 1));
    cf32 *wl = (cf32 *)malloc(sizeof(cf32) * (size_t)nf);
    for (int m = 0; m < nf / 2; ++m) {
        tw[m].re = (float)cos(-2.0 * ORC_PI * m / nf);
        tw[m].im = (float)sin(-2.0 * ORC_PI * m / nf);
    }
    for (int k = 0; k < nf; ++k) {
        wl[k].re = (float)cos(-2.0 * ORC_PI * k / L);
        wl[k].im = (float)sin(-2.0 * ORC_PI * k / L);
    }
    int32_t *pj = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)B);
    orc_pairs(N, pj);
    memset(V, 0, sizeof(float) * 2 * (size_t)B * (size_t)nf);
    if (chunk_frames < 1) chunk_frames = T;
    const int cf = chunk_frames < T ? chunk_frames : T;
    cf32 *Y = (cf32 *)malloc(sizeof(cf32) * (size_t)N * (size_t)cf * (size_t)nf);
    int threads = 1;

    for (int f0 = 0; f0 < T; f0 += cf) {
        const int ft = T - f0 < cf ? T - f0 : cf;
#pragma omp parallel
        {
#ifdef _OPENMP
#pragma omp single
            threads = omp_get_num_threads();
#endif
            float *xf = (float *)malloc(sizeof(float) * (size_t)L);
            cf32 *z = (cf32 *)malloc(sizeof(cf32) * (size_t)nf);
#pragma omp for schedule(static) collapse(2)
            for (int i = 0; i < N; ++i)
                for (int t = 0; t < ft; ++t) {
                    const int8_t *src = x + ((size_t)i * T + (size_t)(f0 + t)) * (size_t)L;
                    const int ok = frame_ok(valid, i, f0 + t, T);
                    for (int n = 0; n < L; ++n) xf[n] = ok ? (float)src[n] : 0.0f;
                    r2c_f32(xf, nf, tw, wl, z, Y + ((size_t)i * cf + t) * (size_t)nf);
                }
            free(z);
            free(xf);
        }
        /* X: per chunk, partial sums over the chunk's frames in frame order, then
           one add into the row (kernels.cu:236-264). Channel blocks of 64 keep
           the partials in cache and vectorise over channels. */
        const int CB = 64;
        const int ncb = (nf + CB - 1) / CB;
#pragma omp parallel for schedule(dynamic) collapse(2)
        for (int b = 0; b < B; ++b)
            for (int cb = 0; cb < ncb; ++cb) {
                const int j = pj[2 * b], k = pj[2 * b + 1];
                const int c0 = cb * CB, c1 = c0 + CB < nf ? c0 + CB : nf;
                float ar[64], ai[64];
                for (int c = 0; c < CB; ++c) { ar[c] = 0.0f; ai[c] = 0.0f; }
                for (int t = 0; t < ft; ++t) {
                    const cf32 *yj = Y + ((size_t)j * cf + t) * (size_t)nf;
                    const cf32 *yk = Y + ((size_t)k * cf + t) * (size_t)nf;
                    for (int c = c0; c < c1; ++c) {
                        const cf32 p = yj[c], q = yk[c];
                        ar[c - c0] += p.re * q.re + p.im * q.im;
                        ai[c - c0] += p.im * q.re - p.re * q.im;
                    }
                }
                float *row = V + (size_t)b * nf * 2;
                for (int c = c0; c < c1; ++c) {
                    row[2 * c] += ar[c - c0];
                    row[2 * c + 1] += ai[c - c0];
                }
            }
    }
    const float s = 1.0f / (float)frames_div;
    for (size_t i = 0; i < 2 * (size_t)B * (size_t)nf; ++i) V[i] *= s;
    free(Y);
    free(pj);
    free(wl);
    free(tw);
    return threads;
}

/* F-engine alone in double: spectra [N][T][nf] complex interleaved. */
void orc_spectra_f64(const int8_t *x, const uint8_t *valid, int N, int T, int L, double *Y) {
    const int nf = L / 2;
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * (size_t)L);
#pragma omp for schedule(dynamic) collapse(2)
        for (int i = 0; i < N; ++i)
            for (int t = 0; t < T; ++t) {
                const int8_t *src = x + ((size_t)i * T + t) * (size_t)L;
                const int ok = frame_ok(valid, i, t, T);
                for (int n = 0; n < L; ++n) buf[n] = ok ? (double)src[n] : 0.0;
                orc_fft_r2c_f64(buf, L, Y + ((size_t)i * T + t) * 2 * (size_t)nf);
            }
        free(buf);
    }
}
