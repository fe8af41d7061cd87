/*
 * corr_cuda.h -- the drop-in C ABI of the MI355X correlator library.
 *
 * This header declares exactly the interface that the reference's GPU library
 * exports (corr-cuda/include/corr_cuda.h, ABI version 7), so that the Rust host
 * (corr-rs/src/ffi.rs:157-219, linked by corr-rs/build.rs:39-57) links against
 * this library unchanged. Every constant, enum value, struct field, field order
 * and function signature below is kept identical on purpose: it is the contract.
 * Only the comments are this library's own. The library behind it is HIP for
 * gfx950: hand-written kernels for the F-engine (mask + int8 unpack + R2C FFT)
 * and the X-engine, with no cuFFT or hipFFT on the path.
 *
 * Model of use (reference corr_cuda.h:14-37):
 *   create                -> accumulator stack zeroed, row 0 active
 *   row_accumulate(seq)*K -> acc[row] += sum over the slot's frames of Y_j conj(Y_k)
 *   row_end(frames)       -> acc[row] *= 1/frames, next row
 *   rows_wait / rows_fetch-> one D2H of every closed row, then clear
 *
 * Staging (reference corr_cuda.h:39-54, 247-294): a pinned ring of block_slots
 * slots per station, frames_per_slot frames of 2*n_freq int8 samples each, then a
 * validity bitmap per slot appended after the station's samples. A clear bit
 * means the frame is missing and it is correlated as zeros.
 *
 * Threading (reference corr_cuda.h:65-71): one driving thread per context;
 * callbacks run on a runtime thread and may only signal.
 */

#ifndef CORR_CUDA_H
#define CORR_CUDA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reference corr_cuda.h:94 -- v7: the validity mask is read by the device and
   cleared by corr_row_slot_free(). */
#define CORR_ABI_VERSION 7u

/* reference corr_cuda.h:113-137 -- how the pair space is split over devices. */
#define CORR_SPLIT_BASELINE_RANGE 0
#define CORR_SPLIT_PAIR_COVERING 1
#define CORR_SPLIT_TIME_SLICE 2

/* reference corr_cuda.h:141 -- size of the device array in corr_config. */
#define CORR_MAX_DEVICES 8u

/* reference corr_cuda.h:143-154. The numeric values are the ABI. On this library
   CORR_ERR_CUDA means "HIP runtime error" and CORR_ERR_CUFFT means "FFT error". */
typedef enum corr_status {
    CORR_OK           = 0,
    CORR_ERR_ARG      = 1,
    CORR_ERR_ABI      = 2,
    CORR_ERR_CUDA     = 3,
    CORR_ERR_CUFFT    = 4,
    CORR_ERR_STATE    = 5,
    CORR_ERR_BUSY     = 6,
    CORR_ERR_NOMEM    = 7,
    CORR_ERR_NOT_IMPL = 8,
    CORR_ERR_INTERNAL = 9
} corr_status;

/* reference corr_cuda.h:171-196. 76 bytes on x86-64; struct_size is checked. */
typedef struct corr_config {
    uint32_t abi_version;     /* must equal CORR_ABI_VERSION                  */
    uint32_t struct_size;     /* must equal sizeof(corr_config)               */
    int32_t  device_id;       /* device ordinal used when n_devices == 0      */
    int32_t  n_station;       /* N inputs; N(N+1)/2 baselines incl. autos     */
    int32_t  n_freq;          /* retained channels; frame = 2*n_freq samples  */
    int32_t  frames_per_slot; /* frames per station in one staging slot       */
    int32_t  chunk_frames;    /* frames per device work unit, <= slot         */
    int32_t  block_slots;     /* staging slots in the ring, >= 2              */
    int32_t  rows_per_fetch;  /* integration rows held before one D2H         */
    int32_t  n_devices;       /* 0 = just device_id, else 1..CORR_MAX_DEVICES */
    int32_t  devices[CORR_MAX_DEVICES];
    int32_t  split_mode;      /* CORR_SPLIT_*                                 */
} corr_config;

/* reference corr_cuda.h:209-219. */
typedef struct corr_device_plan {
    int32_t device_id;
    int32_t n_station_resident;
    int32_t first_baseline;
    int32_t n_baseline;
    int32_t contiguous;
} corr_device_plan;

/* reference corr_cuda.h:229-245. */
typedef struct corr_device_stats {
    int32_t  device_id;
    uint64_t bytes_h2d;
    uint64_t bytes_d2h;
    uint64_t transfers;
    uint64_t ns_accumulate;
    uint64_t ns_slot_free;
    uint64_t ns_row_end;
    uint64_t ns_rows_wait;
    uint64_t ns_fetch;
} corr_device_stats;

/* reference corr_cuda.h:262-294. Frame f of slot s of this station lives at
   base + s*slot_stride + f*frame_stride; the slot's validity bits live at
   base + validity_offset + s*validity_bytes_per_slot, bit f = byte f>>3, bit f&7. */
typedef struct corr_staging {
    int8_t *base;
    size_t  frame_stride;
    size_t  slot_stride;
    size_t  total_bytes;
    size_t  validity_offset;
    size_t  validity_bytes_per_slot;
} corr_staging;

/* reference corr_cuda.h:301 -- version banner; contains "abi=7". */
const char *corr_abi_banner(void);

/* reference corr_cuda.h:307 -- thread-local last failure, never NULL. */
const char *corr_last_error(void);

/* reference corr_cuda.h:317 -- validate, allocate staging + device state.
   *out_ctx is NULL on failure. */
corr_status corr_create(const corr_config *cfg, void **out_ctx);

/* reference corr_cuda.h:320 -- NULL is a no-op. */
corr_status corr_destroy(void *ctx);

/* reference corr_cuda.h:326 */
corr_status corr_staging_info(void *ctx, int32_t station, corr_staging *out);

/* reference corr_cuda.h:356 -- queue slot (slot_seq % block_slots), `frames`
   frames in [1, frames_per_slot], into the current row. */
corr_status corr_row_accumulate(void *ctx, int64_t slot_seq, int32_t frames);

/* reference corr_cuda.h:371 -- block until the slot's H2D is done, clear its
   validity bits, mark it free. */
corr_status corr_row_slot_free(void *ctx, int64_t slot_seq);

/* reference corr_cuda.h:399 */
typedef void (*corr_notify_fn)(void *user);

/* reference corr_cuda.h:401-402, 412 -- the same release in two halves. */
corr_status corr_row_slot_free_submit(void *ctx, int64_t slot_seq,
                                      corr_notify_fn cb, void *user);
corr_status corr_row_slot_free_finish(void *ctx, int64_t slot_seq);

/* reference corr_cuda.h:428 -- close the row, divide by `frames`. */
corr_status corr_row_end(void *ctx, uint64_t frames);

/* reference corr_cuda.h:434, 445, 452, 459 */
corr_status corr_rows_wait(void *ctx);
corr_status corr_rows_wait_submit(void *ctx, corr_notify_fn cb, void *user);
corr_status corr_rows_ready(void *ctx);
corr_status corr_rows_pending(void *ctx, int32_t *out_pending);

/* reference corr_cuda.h:487-488 -- out[(row*n_baseline + b)*n_freq + bin],
   complex64, real first, bin 0 = DC, Nyquist dropped, no FFT normalisation. */
corr_status corr_rows_fetch(void *ctx, void *out, size_t out_bytes,
                            int32_t *out_rows);

/* reference corr_cuda.h:511-512, 520 */
corr_status corr_rows_fetch_submit(void *ctx, void *out, size_t out_bytes,
                                   int32_t *out_rows, corr_notify_fn cb, void *user);
corr_status corr_rows_fetch_ready(void *ctx);

/* reference corr_cuda.h:529, 542 -- baseline b <-> (j,k), j <= k, j outer. */
int32_t corr_n_baseline(int32_t n_station);
corr_status corr_baseline_pairs(int32_t n_station, int32_t *out_jk, int32_t out_len);

/* reference corr_cuda.h:561-584 -- what the device split was. */
int32_t corr_n_devices(void *ctx);
corr_status corr_device_plan_get(void *ctx, int32_t device_index,
                                 corr_device_plan *out);
corr_status corr_device_stations(void *ctx, int32_t device_index,
                                 int32_t *out_stations, int32_t out_len);
corr_status corr_device_baselines(void *ctx, int32_t device_index,
                                  int32_t *out_baselines, int32_t out_len);
corr_status corr_device_stats_get(void *ctx, int32_t device_index,
                                  corr_device_stats *out);

/* reference corr_cuda.h:603-604 -- pinned host memory for fetch buffers. */
corr_status corr_host_alloc(size_t bytes, void **out);
void corr_host_free(void *p);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* CORR_CUDA_H */
