/* abi_check.c -- the drop-in ABI driven from C, linked against libcorr_cuda.
 *
 * The reference keeps its ABI contract in C programs linked against the
 * library (corr-cuda/tests/abi_smoke.c, async_notify.c, golden.c, registered at
 * corr-cuda/CMakeLists.txt:84-163), so that a fault can be placed on the
 * library's side or the caller's. This program does the same for this library,
 * written against include/corr_cuda.h alone:
 *
 *   abi_check          link check, no device work: the struct sizes of the
 *                      x86-64 ABI are asserted at compile time, all 26 entry
 *                      points are referenced (so the link fails if one is
 *                      missing), the banner is printed and must say abi=7.
 *   abi_check run      the contract on device 0: numbering, config rejections,
 *                      staging geometry, slot and row state rules, the
 *                      asynchronous notifications, device statistics, and two
 *                      closed-form answers through the whole pipeline (zero in,
 *                      zero out; one impulse per station and frame, whose
 *                      visibilities are a_j a_k e^{-2 pi i c (p_j - p_k) / L}
 *                      summed over the present frames and divided by n_cum).
 *
 * Exit status 0 when every check holds; every failed check prints its line.
 */
#define _POSIX_C_SOURCE 200809L
#include "corr_cuda.h"

#include <math.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

_Static_assert(sizeof(corr_config) == 76, "corr_config is 76 bytes on x86-64");
_Static_assert(sizeof(corr_device_plan) == 20, "corr_device_plan is 20 bytes");
_Static_assert(sizeof(corr_device_stats) == 72, "corr_device_stats is 72 bytes");
_Static_assert(sizeof(corr_staging) == 48, "corr_staging is 48 bytes");
_Static_assert(CORR_ABI_VERSION == 7u, "ABI version 7");
_Static_assert(CORR_MAX_DEVICES == 8u, "eight device slots");
_Static_assert(CORR_ERR_CUDA == 3 && CORR_ERR_CUFFT == 4 && CORR_ERR_INTERNAL == 9, "status values");

/* Every entry point of the ABI, by address: a missing export fails the link. */
static void *const k_entry_points[] = {
    (void *)corr_abi_banner,        (void *)corr_last_error,        (void *)corr_create,
    (void *)corr_destroy,           (void *)corr_staging_info,      (void *)corr_row_accumulate,
    (void *)corr_row_slot_free,     (void *)corr_row_slot_free_submit,
    (void *)corr_row_slot_free_finish,
    (void *)corr_row_end,           (void *)corr_rows_wait,         (void *)corr_rows_wait_submit,
    (void *)corr_rows_ready,        (void *)corr_rows_pending,      (void *)corr_rows_fetch,
    (void *)corr_rows_fetch_submit, (void *)corr_rows_fetch_ready,  (void *)corr_n_baseline,
    (void *)corr_baseline_pairs,    (void *)corr_n_devices,         (void *)corr_device_plan_get,
    (void *)corr_device_stations,   (void *)corr_device_baselines,  (void *)corr_device_stats_get,
    (void *)corr_host_alloc,        (void *)corr_host_free,
};
#define N_ENTRY (int)(sizeof k_entry_points / sizeof k_entry_points[0])
_Static_assert(sizeof k_entry_points / sizeof k_entry_points[0] == 26, "26 entry points");

static int g_failed = 0;

#define EXPECT(cond, ...)                                                      \
    do {                                                                       \
        if (!(cond)) {                                                         \
            ++g_failed;                                                        \
            printf("  failed at line %d: ", __LINE__);                         \
            printf(__VA_ARGS__);                                               \
            printf(" [last error: %s]\n", corr_last_error());                  \
        }                                                                      \
    } while (0)

#define EXPECT_STATUS(call, want)                                              \
    do {                                                                       \
        corr_status got_ = (call);                                             \
        EXPECT(got_ == (want), "%s returned %d, expected %d", #call, (int)got_, \
               (int)(want));                                                   \
    } while (0)

static corr_config small_config(void) {
    corr_config c;
    memset(&c, 0, sizeof c);
    c.abi_version = CORR_ABI_VERSION;
    c.struct_size = (uint32_t)sizeof c;
    c.n_station = 3;
    c.n_freq = 2048;
    c.frames_per_slot = 12;
    c.chunk_frames = 8; /* slots of 12 frames: one chunk of 8, one of 4 */
    c.block_slots = 3;
    c.rows_per_fetch = 2;
    return c;
}

static int link_check(void) {
    const char *b = corr_abi_banner();
    printf("banner: %s\n", b ? b : "(null)");
    int missing = 0;
    for (int i = 0; i < N_ENTRY; ++i) missing += k_entry_points[i] == NULL;
    EXPECT(b != NULL && strstr(b, "abi=7") != NULL, "banner lacks abi=7");
    EXPECT(missing == 0, "%d entry points resolved to NULL", missing);
    EXPECT(corr_n_baseline(24) == 300, "corr_n_baseline(24)");
    printf("linked %d entry points\n", N_ENTRY);
    return g_failed;
}

static void check_numbering(void) {
    EXPECT(corr_n_baseline(0) == -1 && corr_n_baseline(1) == 1 && corr_n_baseline(40) == 820,
           "baseline counts");
    int32_t jk[2 * 6];
    EXPECT_STATUS(corr_baseline_pairs(3, jk, 12), CORR_OK);
    const int32_t want[12] = {0, 0, 0, 1, 0, 2, 1, 1, 1, 2, 2, 2};
    EXPECT(memcmp(jk, want, sizeof want) == 0, "pair order of 3 stations");
    EXPECT_STATUS(corr_baseline_pairs(3, jk, 11), CORR_ERR_ARG);
    EXPECT_STATUS(corr_baseline_pairs(0, jk, 0), CORR_ERR_ARG);
}

static void check_rejections(void) {
    void *ctx = (void *)&ctx;
    corr_config c;
    EXPECT_STATUS(corr_create(NULL, &ctx), CORR_ERR_ARG);
    EXPECT(ctx == NULL, "out_ctx not cleared on failure");
    EXPECT(strlen(corr_last_error()) > 0, "a failure left no message");
    c = small_config();
    EXPECT_STATUS(corr_create(&c, NULL), CORR_ERR_ARG);
    c.abi_version = 6;
    EXPECT_STATUS(corr_create(&c, &ctx), CORR_ERR_ABI);
    c = small_config();
    c.struct_size = 72;
    EXPECT_STATUS(corr_create(&c, &ctx), CORR_ERR_ABI);
    c = small_config();
    c.n_freq = 3072;
    EXPECT_STATUS(corr_create(&c, &ctx), CORR_ERR_ARG);
    c = small_config();
    c.chunk_frames = 13;
    EXPECT_STATUS(corr_create(&c, &ctx), CORR_ERR_ARG);
    c = small_config();
    c.block_slots = 1;
    EXPECT_STATUS(corr_create(&c, &ctx), CORR_ERR_ARG);
    c = small_config();
    c.rows_per_fetch = 0;
    EXPECT_STATUS(corr_create(&c, &ctx), CORR_ERR_ARG);
    c = small_config();
    c.split_mode = 3;
    EXPECT_STATUS(corr_create(&c, &ctx), CORR_ERR_ARG);
    c = small_config();
    c.n_devices = 9;
    EXPECT_STATUS(corr_create(&c, &ctx), CORR_ERR_ARG);
    c = small_config();
    c.n_devices = 2; /* a repeated ordinal is refused before the device count matters */
    c.devices[0] = 0;
    c.devices[1] = 0;
    EXPECT_STATUS(corr_create(&c, &ctx), CORR_ERR_ARG);
    EXPECT(ctx == NULL, "out_ctx not cleared on failure");
    EXPECT_STATUS(corr_destroy(NULL), CORR_OK);
}

/* Counting callback: may only signal (no ABI calls from the runtime thread). */
static void count_cb(void *user) { atomic_fetch_add((atomic_int *)user, 1); }

static int wait_count(atomic_int *n, int want) {
    for (int i = 0; i < 20000; ++i) { /* up to 20 s */
        if (atomic_load(n) >= want) return 1;
        struct timespec ts = {0, 1000000};
        nanosleep(&ts, NULL);
    }
    return 0;
}

/* Fill slot `slot` of every station: frame f of station s holds one impulse of
   amplitude amp[s] at sample pos[s] (or all zeros when amp is NULL); `present`
   marks which frames get their validity bit. */
static void fill_slot(void *ctx, const corr_config *c, int slot, int frames, const int *amp,
                      const int *pos, const unsigned char *present) {
    for (int s = 0; s < c->n_station; ++s) {
        corr_staging st;
        EXPECT_STATUS(corr_staging_info(ctx, s, &st), CORR_OK);
        for (int f = 0; f < frames; ++f) {
            int8_t *fr = st.base + (size_t)slot * st.slot_stride + (size_t)f * st.frame_stride;
            memset(fr, 0, st.frame_stride);
            if (amp != NULL) fr[pos[s]] = (int8_t)amp[s];
            if (present[f]) {
                uint8_t *m = (uint8_t *)st.base + st.validity_offset +
                             (size_t)slot * st.validity_bytes_per_slot;
                m[f >> 3] |= (uint8_t)(1u << (f & 7));
            }
        }
    }
}

static void check_pipeline(void) {
    const corr_config c = small_config();
    const int N = c.n_station, nf = c.n_freq, L = 2 * nf, B = N * (N + 1) / 2;
    void *ctx = NULL;
    EXPECT_STATUS(corr_create(&c, &ctx), CORR_OK);
    if (ctx == NULL) return;
    EXPECT(corr_n_devices(ctx) == 1, "one device");

    /* staging geometry (reference corr_cuda.h:262-294) */
    corr_staging st;
    EXPECT_STATUS(corr_staging_info(ctx, 0, &st), CORR_OK);
    EXPECT(st.frame_stride == (size_t)L && st.slot_stride == (size_t)L * 12 &&
               st.validity_offset == (size_t)L * 12 * 3 && st.validity_bytes_per_slot == 2 &&
               st.total_bytes == (size_t)L * 12 * 3 + 2 * 3,
           "staging strides");
    corr_staging st1;
    EXPECT_STATUS(corr_staging_info(ctx, 1, &st1), CORR_OK);
    EXPECT(st1.base == st.base + st.total_bytes, "stations follow each other");
    EXPECT_STATUS(corr_staging_info(ctx, N, &st1), CORR_ERR_ARG);

    /* row 0: zeros in, zeros out; row 1: impulses, frames 1 and 6 of the first
       slot missing (garbage-free zero frames are not needed: missing frames are
       never read) */
    static const int amp[3] = {3, -5, 7}, pos[3] = {0, 1, 6};
    unsigned char all[12], some[12];
    for (int f = 0; f < 12; ++f) {
        all[f] = 1;
        some[f] = (unsigned char)(f != 1 && f != 6);
    }
    int64_t seq = 0;
    fill_slot(ctx, &c, 0, 12, NULL, pos, all);
    EXPECT_STATUS(corr_row_accumulate(ctx, seq, 12), CORR_OK);
    EXPECT_STATUS(corr_row_accumulate(ctx, seq + 3, 12), CORR_ERR_BUSY); /* same physical slot */
    EXPECT_STATUS(corr_row_slot_free(ctx, seq + 1), CORR_ERR_STATE);     /* never submitted */
    EXPECT_STATUS(corr_row_accumulate(ctx, -1, 12), CORR_ERR_ARG);
    EXPECT_STATUS(corr_row_accumulate(ctx, seq + 1, 13), CORR_ERR_ARG);
    EXPECT_STATUS(corr_row_end(ctx, 12), CORR_OK);
    ++seq;

    fill_slot(ctx, &c, 1, 12, amp, pos, some);
    EXPECT_STATUS(corr_row_accumulate(ctx, seq, 12), CORR_OK);
    ++seq;
    fill_slot(ctx, &c, 2, 5, amp, pos, all); /* a partial slot of 5 frames */
    EXPECT_STATUS(corr_row_accumulate(ctx, seq, 5), CORR_OK);
    ++seq;
    const uint64_t n_cum = 20; /* nominal count: divides the 15 present frames */
    EXPECT_STATUS(corr_row_end(ctx, n_cum), CORR_OK);
    EXPECT_STATUS(corr_row_end(ctx, n_cum), CORR_ERR_BUSY); /* both rows full */
    EXPECT_STATUS(corr_row_end(ctx, 0), CORR_ERR_ARG);

    /* asynchronous releases: one callback each, then finish */
    atomic_int freed = 0;
    for (int64_t s = 0; s < seq; ++s)
        EXPECT_STATUS(corr_row_slot_free_submit(ctx, s, count_cb, &freed), CORR_OK);
    EXPECT(wait_count(&freed, (int)seq), "slot callbacks: %d of %d", atomic_load(&freed), (int)seq);
    for (int64_t s = 0; s < seq; ++s) EXPECT_STATUS(corr_row_slot_free_finish(ctx, s), CORR_OK);
    EXPECT_STATUS(corr_row_slot_free_finish(ctx, 0), CORR_ERR_STATE);
    /* a released slot has its validity bits cleared */
    {
        const uint8_t *m = (const uint8_t *)st.base + st.validity_offset;
        EXPECT(m[0] == 0 && m[1] == 0, "validity bits of a released slot");
    }

    atomic_int rows_cb = 0;
    EXPECT_STATUS(corr_rows_wait_submit(ctx, count_cb, &rows_cb), CORR_OK);
    EXPECT(wait_count(&rows_cb, 1), "row callback");
    EXPECT_STATUS(corr_rows_ready(ctx), CORR_OK);
    int32_t pending = -1;
    EXPECT_STATUS(corr_rows_pending(ctx, &pending), CORR_OK);
    EXPECT(pending == 2, "pending rows %d", pending);

    const size_t out_bytes = (size_t)2 * B * nf * 8;
    float *out = NULL;
    EXPECT_STATUS(corr_host_alloc(out_bytes, (void **)&out), CORR_OK);
    if (out == NULL) {
        corr_destroy(ctx);
        return;
    }
    int32_t rows = -1;
    EXPECT_STATUS(corr_rows_fetch(ctx, out, out_bytes - 8, &rows), CORR_ERR_ARG);
    atomic_int fetched = 0;
    EXPECT_STATUS(corr_rows_fetch_submit(ctx, out, out_bytes, &rows, count_cb, &fetched), CORR_OK);
    EXPECT(rows == 2, "rows known at submit: %d", rows);
    EXPECT(wait_count(&fetched, 1), "fetch callback");
    EXPECT_STATUS(corr_rows_fetch_ready(ctx), CORR_OK);
    EXPECT_STATUS(corr_rows_pending(ctx, &pending), CORR_OK);
    EXPECT(pending == 0, "rows still pending after the fetch");

    /* row 0 exactly zero */
    size_t nonzero = 0;
    for (size_t i = 0; i < (size_t)B * nf * 2; ++i) nonzero += out[i] != 0.0f;
    EXPECT(nonzero == 0, "zero input gave %zu nonzero values", nonzero);

    /* row 1: V_jk[c] = (15/20) a_j a_k exp(-2 pi i c (p_j - p_k) / L) */
    double num = 0.0, den = 0.0;
    const float *row1 = out + (size_t)B * nf * 2;
    int b = 0;
    for (int j = 0; j < N; ++j)
        for (int k = j; k < N; ++k, ++b)
            for (int ch = 0; ch < nf; ++ch) {
                const double ph = -2.0 * 3.14159265358979323846 * (double)ch * (double)(pos[j] - pos[k]) / (double)L;
                const double mag = 15.0 / (double)n_cum * amp[j] * amp[k];
                const double re = mag * cos(ph), im = mag * sin(ph);
                const float *v = row1 + ((size_t)b * nf + ch) * 2;
                num += (v[0] - re) * (v[0] - re) + (v[1] - im) * (v[1] - im);
                den += re * re + im * im;
            }
    const double rel = sqrt(num / den);
    printf("impulse visibilities: relative RMS vs closed form %.3e\n", rel);
    EXPECT(rel < 1e-5, "impulse visibilities off by %.3e", rel);

    /* an empty fetch returns zero rows and arms no callback */
    atomic_int none = 0;
    EXPECT_STATUS(corr_rows_fetch_submit(ctx, out, out_bytes, &rows, count_cb, &none), CORR_OK);
    EXPECT(rows == 0, "empty fetch gave %d rows", rows);
    EXPECT_STATUS(corr_rows_fetch(ctx, out, out_bytes, &rows), CORR_OK);

    /* device plan, stations, baselines, statistics of the one device */
    corr_device_plan plan;
    EXPECT_STATUS(corr_device_plan_get(ctx, 0, &plan), CORR_OK);
    EXPECT(plan.n_station_resident == N && plan.first_baseline == 0 && plan.n_baseline == B &&
               plan.contiguous == 1,
           "device plan");
    int32_t stations[3], baselines[6];
    EXPECT_STATUS(corr_device_stations(ctx, 0, stations, 3), CORR_OK);
    EXPECT(stations[0] == 0 && stations[2] == 2, "resident stations");
    EXPECT_STATUS(corr_device_baselines(ctx, 0, baselines, 6), CORR_OK);
    EXPECT(baselines[0] == 0 && baselines[5] == 5, "device baselines");
    EXPECT_STATUS(corr_device_plan_get(ctx, 1, &plan), CORR_ERR_ARG);
    corr_device_stats ds;
    EXPECT_STATUS(corr_device_stats_get(ctx, 0, &ds), CORR_OK);
    EXPECT(ds.bytes_d2h == out_bytes && ds.bytes_h2d > 0 && ds.transfers > 0, "device stats");

    EXPECT(wait_count(&none, 0) && atomic_load(&none) == 0, "a callback fired for an empty fetch");
    corr_host_free(out);
    EXPECT_STATUS(corr_destroy(ctx), CORR_OK);
}

int main(int argc, char **argv) {
    if (argc < 2) return link_check() ? 1 : 0;
    if (strcmp(argv[1], "run") != 0) {
        fprintf(stderr, "usage: %s [run]\n", argv[0]);
        return 2;
    }
    link_check();
    printf("numbering\n");
    check_numbering();
    printf("config rejections\n");
    check_rejections();
    printf("pipeline on device 0\n");
    check_pipeline();
    printf(g_failed ? "FAILED: %d checks\n" : "all checks passed\n", g_failed);
    return g_failed ? 1 : 0;
}
