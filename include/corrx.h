/*
 * corrx.h -- engine-level entry points of the MI355X correlator library.
 *
 * The drop-in ABI (corr_cuda.h) takes frames through a pinned staging ring and
 * returns rows through the host. These calls expose the same device stages on
 * device-resident buffers, launched on a caller-supplied HIP stream, for
 * callers that already hold their samples in HBM: the benchmark, the parity
 * tests, and geometries the drop-in ABI refuses by its packet rule
 * (n_freq % 2048 != 0, e.g. n_freq = 1024). They have no counterpart in the
 * reference's ABI; the work they run is the reference's per-chunk device work
 * (corr-cuda/src/kernels.cu:810-845) and row scale (kernels.cu:878-903).
 *
 * Plain pointers and sizes only. `stream` is a hipStream_t passed as void*
 * (NULL = the null stream). Device pointers must belong to the device that was
 * current when the engine was created; that device must be current when a stage
 * is called.
 */
#ifndef CORRX_H
#define CORRX_H

#include <stddef.h>
#include <stdint.h>

#include "corr_cuda.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Spectrum representation between the F- and the X-engine. */
#define CORRX_FORM_FP32 0 /* complex64 spectra, FP32 X-engine: the drop-in numerics */
#define CORRX_FORM_FP16 1 /* FP16 spectra, per-(input,channel) gains, FP16 MFMA     */
#define CORRX_FORM_INT8 2 /* int8 spectra in [-7,7], gains, int8 MFMA, exact int32 */

typedef struct corrx_shape {
    int32_t n_station;    /* inputs, >= 1                                    */
    int32_t n_freq;       /* retained channels: 2^k in [512, 16384], or a multiple of 2048 up to 1048576 */
    int32_t chunk_frames; /* most frames one fengine/xengine call handles;
                             chunk_frames * n_freq < 2^32                 */
    int32_t form;         /* CORRX_FORM_*                                    */
} corrx_shape;

/* Create an engine on the current device: FP64-built twiddle tables, baseline
   tables, a spectrum buffer of chunk_frames frames. */
corr_status corrx_engine_create(const corrx_shape *shape, void **out_engine);
corr_status corrx_engine_destroy(void *engine);

/* F-engine: mask + int8 -> f32 + R2C FFT of `frames` frames of every station.
   Frame f of station s is at d_in + s*station_pitch + f*2*n_freq (bytes).
   d_mask may be NULL (every frame present); otherwise frame f of station s is
   present when bit (mask_bit0 + f) of d_mask + s*mask_pitch is set, bit b
   being (byte b>>3 >> (b&7)) & 1. A missing frame is transformed as zeros. */
corr_status corrx_fengine(void *engine, const int8_t *d_in, size_t station_pitch,
                          const uint8_t *d_mask, size_t mask_pitch, int32_t mask_bit0,
                          int32_t frames, void *stream);

/* X-engine: d_acc[b*n_freq + c] += sum over the frames of the last fengine call
   of Y_j conj(Y_k), baseline b = (j,k) in corr_baseline_pairs order. The
   integrator element is the form's (corrx_integrator_bytes):
     FP32: complex64
     FP16: 4 floats (sum re, sum im, compensation re, compensation im); the sum
           of a call is added with Neumaier's compensated summation
     INT8: 2 int32 (re, im), exact; at most 2^31/98 frames per integration
   A zeroed integrator starts an integration. */
corr_status corrx_xengine(void *engine, void *d_acc, int32_t frames, void *stream);

/* F- and X-engine over `frames` frames (any count; frame f of station s at
   d_in + s*station_pitch + f*2*n_freq, mask bits as corrx_fengine), added into
   d_acc as corrx_xengine adds: the same bytes as corrx_fengine + corrx_xengine
   over successive runs of at most chunk_frames frames. In the FP16 and INT8
   forms with more than one run, the X-engine runs on the engine's own stream
   over a second spectrum buffer, so that it overlaps the next run's F-engine;
   the FP32 form (whose X-engine is VALU-bound like the F-engine) and single
   runs go in turn on `stream`. Everything is ordered after the work already
   enqueued on `stream`, and `stream` waits for the last X-engine before the
   call returns. The first overlapped call allocates the second spectrum buffer
   (as large as the first).
   Replaces the per-chunk loop of reference kernels.cu:810-845 for callers that
   hold a whole integration in device memory. */
corr_status corrx_integrate(void *engine, const int8_t *d_in, size_t station_pitch,
                            const uint8_t *d_mask, size_t mask_pitch, int32_t mask_bit0,
                            int64_t frames, void *d_acc, void *stream);

/* Batches: one dump of several independent sub-arrays (engines of any shapes
   and forms) at once. No counterpart in the reference, which runs one array per
   context; this is the batched form of corrx_integrate + corrx_dump.

   Per engine, a run does, on a branch stream of its own forked from `stream`:
   zero d_acc (flag CORRX_BATCH_ZERO; a fill kernel), corrx_integrate over
   `frames` frames, and corrx_dump of d_acc as `dump_frames` frames into d_out
   (flag CORRX_BATCH_DUMP); `stream` waits for every branch. With use_graph != 0
   the run is captured into a hipGraph on first use and replayed while the
   items stay byte-identical (other items: a new capture). The engines must be
   distinct, on the current device, and outlive the batch. */
#define CORRX_BATCH_ZERO 1
#define CORRX_BATCH_DUMP 2

typedef struct corrx_batch_item {
    const int8_t *d_in;      /* as corrx_integrate */
    size_t station_pitch;
    const uint8_t *d_mask;   /* NULL: every frame present */
    size_t mask_pitch;
    int32_t mask_bit0;
    int32_t flags;           /* CORRX_BATCH_ZERO | CORRX_BATCH_DUMP */
    int64_t frames;          /* frames integrated in this run (>= 0) */
    void *d_acc;             /* the engine's integrator, [B][n_freq] of its element */
    uint64_t dump_frames;    /* nominal frame count of the dump (> 0 with CORRX_BATCH_DUMP) */
    void *d_out;             /* complex64 [B][n_freq] (with CORRX_BATCH_DUMP) */
} corrx_batch_item;          /* 72 bytes on x86-64 */

corr_status corrx_batch_create(void *const *engines, int32_t n_engines, void **out_batch);
corr_status corrx_batch_destroy(void *batch);
corr_status corrx_batch_run(void *batch, const corrx_batch_item *items, int32_t use_graph, void *stream);

/* Bytes of one integrator element (per baseline and channel) of the engine's form. */
corr_status corrx_integrator_bytes(void *engine, size_t *bytes);

/* Visibilities of an integration of `frames` frames (the caller's nominal count,
   as corr_row_end): d_out[b*n_freq + c] complex64, the dump layout of
   corr_rows_fetch. FP32: d_acc * (1.0f/(float)frames) (d_out may equal d_acc);
   FP16/INT8: integrator / (g_j[c] g_k[c] frames), evaluated in FP64.
   INT8: CORR_ERR_ARG for frames > 2^31/98 = 21913098, the most its int32
   integrator holds without wrapping. */
corr_status corrx_dump(void *engine, const void *d_acc, uint64_t frames, void *d_out, void *stream);

/* FP16/INT8 forms: set the per-(input, channel) gains from the mean power of
   `frames` frames (arguments as corrx_fengine, any frame count; missing frames
   are not counted). FP16 gain = 2^(8-E) for a component RMS m*2^E, m in
   [0.5,1); INT8 gain = 1/(0.35*RMS). A channel with no power keeps gain 1.
   Overwrites the spectrum buffer. */
corr_status corrx_calibrate(void *engine, const int8_t *d_in, size_t station_pitch,
                            const uint8_t *d_mask, size_t mask_pitch, int32_t mask_bit0,
                            int32_t frames, void *stream);

/* FP16/INT8 forms: gains [station][n_freq] (float, positive, finite) copied from
   device memory; and the engine's current gain table (device pointer). */
corr_status corrx_set_gains(void *engine, const float *d_gains, void *stream);
corr_status corrx_gains(void *engine, const float **d_gains);

/* Matrix-core X-engine (every form up to 48 inputs): channels per X-engine
   workgroup (16 or 32; 32 only up to 24 inputs) and the channel tile of the
   spectrum buffer's order (FP16/INT8: that width, or n_freq for the plain
   order; FP32: n_freq only). The engine picks both at creation by timing
   F- and X-engine with every combination that fits its shape (cached per
   process, form, input count, n_freq, launch size and device); set overrides
   the choice (before the next corrx_fengine call), get reports it. The
   environment variable CORRX_XTILE=<channels>,<spectra_tile> (spectra_tile 0 =
   n_freq) pins the choice at creation, without timing, for every engine whose
   shape it fits. The FP32 form also has a vector X-engine (FMA on the VALU):
   channels 0 with spectra_tile n_freq selects it (CORRX_XTILE=0,0 at creation);
   an FP32 engine of more than 48 inputs always runs it, and set fails there
   for any other choice. get reports channels 0 for it. */
corr_status corrx_set_xtile(void *engine, int32_t channels, int32_t spectra_tile);
corr_status corrx_get_xtile(void *engine, int32_t *channels, int32_t *spectra_tile);

/* Matrix-core X-engine: time-K, the frames staged per step, 1, 2 or 4 times the
   K extent of the form's MFMA (FP32 16, FP16 32, int8 64 frames). Deeper steps
   need more LDS and registers per workgroup and are available where the
   staging fits (n_station * (frames per step / K extent) <= 64). The engine's
   tuner picks it together with the tile (CORRX_XTILE's optional third field,
   <channels>,<spectra_tile>,<frames_per_step>, pins it); set overrides it (after
   corrx_set_xtile, which falls back to one K extent where the new width cannot
   hold the depth), get reports it (0 for the FP32 vector X-engine). Results do
   not depend on it beyond the order of FP32 additions. No counterpart in the
   reference (its kStagedFrames, corr-cuda/src/kernels.cu:103-106, is fixed). */
corr_status corrx_set_xstep(void *engine, int32_t frames_per_step);
corr_status corrx_get_xstep(void *engine, int32_t *frames_per_step);

/* One FP32-form integration that allocates nothing: no counterpart in the
   reference; it is for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   The work is corrx_integrate + corrx_dump on an FP32 engine: F-engine,
   matrix-core X-engine (16 channels per workgroup, spectra in tiles of 16
   channels) and the row scale, over `frames` frames (arguments as
   corrx_integrate), into d_out: complex64 [B][n_freq], zeroed by the call,
   accumulated, then multiplied by 1.0f/(float)dump_frames (as corr_row_end).
   In place of an engine's buffers it uses a static scratch array of 256 MiB
   that each GPU the library runs on holds until the process ends: twiddle
   tables, then runs of K frames of validity flags and spectra, with
   K = frames if that fits, else floor((256 MiB - 2 * align256(8 * n_freq)
   - 512) / (n_station * (8 * n_freq + 4))) (n_freq a power of two; 85 frames
   at 24 inputs and n_freq 16384). The visibilities are byte-identical to an
   FP32 engine of chunk_frames K with corrx_set_xtile(e, 16, 16).
   Limits: 1 <= n_station <= 48; n_freq = 2^k in [512, 16384], 6144, 10240,
   12288 or 14336 (CORR_ERR_NOT_IMPL otherwise); a frame that does not fit the
   scratch is refused with CORR_ERR_ARG. Every call on one GPU shares that GPU's
   array and leaves its contents behind: calls that overlap on one GPU (other
   streams, other threads) race. The first call on a device looks the array's
   address up (hipGetSymbolAddress) and raises the X-engine's LDS limit; it
   allocates nothing either. All launches go on `stream`, with no memset node
   and no host synchronisation, so a call can be captured into a graph. */
corr_status corrx_correlate_noalloc(int32_t n_station, int32_t n_freq, const int8_t *d_in, size_t station_pitch,
                                    const uint8_t *d_mask, size_t mask_pitch, int32_t mask_bit0, int64_t frames,
                                    uint64_t dump_frames, void *d_out, void *stream);

/* d_acc *= 1.0f/(float)frames over n_complex values. */
corr_status corrx_scale(void *d_acc, size_t n_complex, uint64_t frames, void *stream);

/* The engine's spectrum buffer, every form: [station][n_freq/T][chunk_frames][T]
   with T the spectra tile of corrx_get_xtile (T = n_freq is the plain
   [station][frame][channel] order), of complex64 (FP32), two IEEE halves in 32
   bits, real in the low half (FP16), or two int8 in 16 bits, real in the low
   byte (INT8); the FP16 and INT8 values are the spectrum times the gain of
   their (input, channel). station_stride_elems = chunk_frames * n_freq. */
corr_status corrx_spectra(void *engine, void **d_spectra, size_t *station_stride_elems);

/* A drop-in context (corr_cuda.h) whose staging holds samples of `sample_bits`
   bits: 8 is exactly corr_create; 4 packs two 4-bit two's complement samples
   per byte, sample 2m of a frame in bits 0-3 and sample 2m+1 in bits 4-7 of
   byte m, values -8..7. With 4 bits the staging's frame_stride is n_freq bytes
   (corr_staging_info reports the packed geometry) and each chunk's
   host-to-device copy moves half the bytes. On the GPU the chunk is unpacked to
   int8 and transformed by the F-engine of corr_create, so the rows are
   byte-identical to corr_create's on the same samples; every other call
   behaves as for corr_create. The caller packs and clips; values outside -8..7
   cannot be represented.
   No counterpart in the reference's ABI: it is the reference's proposed next
   step "4-bit staging packing" (HANDOFF.md:1437-1439, 1863-1864), for receive
   paths whose samples fit 4 bits and whose host link bounds the rate.
   Errors as corr_create, plus CORR_ERR_ARG for sample_bits other than 8 or 4. */
corr_status corrx_create_packed(const corr_config *cfg, int32_t sample_bits, void **out_ctx);

#ifdef __cplusplus
}
#endif

#endif /* CORRX_H */
