// corrx_abi.cpp -- C entry points of the device engine (include/corrx.h).
#include <hip/hip_runtime.h>

#include <vector>

#include "batch.hpp"
#include "corrx.h"
#include "engine.hpp"

using corr::Batch;
using corr::Engine;
using corr::fail;

static Engine *as_engine(void *e) { return static_cast<Engine *>(e); }
static hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }

extern "C" corr_status corrx_engine_create(const corrx_shape *shape, void **out_engine) {
    if (out_engine == nullptr) return fail(CORR_ERR_ARG, "out_engine is NULL");
    *out_engine = nullptr;
    if (shape == nullptr) return fail(CORR_ERR_ARG, "shape is NULL");
    if (shape->form < CORRX_FORM_FP32 || shape->form > CORRX_FORM_INT8)
        return corr::failf(CORR_ERR_ARG, "form=%d is not a CORRX_FORM_*", shape->form);
    corr::EngineShape s;
    s.n_station = shape->n_station;
    s.n_freq = shape->n_freq;
    s.chunk_frames = shape->chunk_frames;
    s.form = static_cast<corr::Form>(shape->form);
    Engine *e = nullptr;
    corr_status st = Engine::create(s, &e);
    if (st != CORR_OK) return st;
    *out_engine = e;
    return CORR_OK;
}

extern "C" corr_status corrx_engine_destroy(void *engine) {
    delete as_engine(engine);
    return CORR_OK;
}

extern "C" corr_status corrx_fengine(void *engine, const int8_t *d_in, size_t station_pitch,
                                     const uint8_t *d_mask, size_t mask_pitch, int32_t mask_bit0,
                                     int32_t frames, void *stream) {
    if (engine == nullptr) return fail(CORR_ERR_ARG, "engine is NULL");
    return as_engine(engine)->fengine(d_in, station_pitch, d_mask, mask_pitch, mask_bit0, frames,
                                      as_stream(stream));
}

extern "C" corr_status corrx_xengine(void *engine, void *d_acc, int32_t frames, void *stream) {
    if (engine == nullptr) return fail(CORR_ERR_ARG, "engine is NULL");
    return as_engine(engine)->xengine(d_acc, frames, as_stream(stream));
}

extern "C" corr_status corrx_integrate(void *engine, const int8_t *d_in, size_t station_pitch,
                                       const uint8_t *d_mask, size_t mask_pitch, int32_t mask_bit0,
                                       int64_t frames, void *d_acc, void *stream) {
    if (engine == nullptr) return fail(CORR_ERR_ARG, "engine is NULL");
    return as_engine(engine)->integrate(d_in, station_pitch, d_mask, mask_pitch, mask_bit0, frames, d_acc,
                                        as_stream(stream));
}

extern "C" corr_status corrx_integrator_bytes(void *engine, size_t *bytes) {
    if (engine == nullptr) return fail(CORR_ERR_ARG, "engine is NULL");
    if (bytes == nullptr) return fail(CORR_ERR_ARG, "bytes is NULL");
    *bytes = as_engine(engine)->integrator_elem_bytes();
    return CORR_OK;
}

extern "C" corr_status corrx_dump(void *engine, const void *d_acc, uint64_t frames, void *d_out,
                                  void *stream) {
    if (engine == nullptr) return fail(CORR_ERR_ARG, "engine is NULL");
    return as_engine(engine)->dump(d_acc, frames, static_cast<float2 *>(d_out), as_stream(stream));
}

extern "C" corr_status corrx_calibrate(void *engine, const int8_t *d_in, size_t station_pitch,
                                       const uint8_t *d_mask, size_t mask_pitch, int32_t mask_bit0,
                                       int32_t frames, void *stream) {
    if (engine == nullptr) return fail(CORR_ERR_ARG, "engine is NULL");
    return as_engine(engine)->calibrate(d_in, station_pitch, d_mask, mask_pitch, mask_bit0, frames,
                                        as_stream(stream));
}

extern "C" corr_status corrx_set_gains(void *engine, const float *d_gains, void *stream) {
    if (engine == nullptr) return fail(CORR_ERR_ARG, "engine is NULL");
    return as_engine(engine)->set_gains(d_gains, as_stream(stream));
}

extern "C" corr_status corrx_gains(void *engine, const float **d_gains) {
    if (engine == nullptr) return fail(CORR_ERR_ARG, "engine is NULL");
    if (d_gains == nullptr) return fail(CORR_ERR_ARG, "output pointer is NULL");
    if (as_engine(engine)->gains() == nullptr) return fail(CORR_ERR_ARG, "the FP32 form has no gains");
    *d_gains = as_engine(engine)->gains();
    return CORR_OK;
}

extern "C" corr_status corrx_scale(void *d_acc, size_t n_complex, uint64_t frames, void *stream) {
    if (d_acc == nullptr) return fail(CORR_ERR_ARG, "d_acc is NULL");
    return Engine::scale(static_cast<float2 *>(d_acc), n_complex, frames, as_stream(stream));
}

extern "C" corr_status corrx_spectra(void *engine, void **d_spectra, size_t *station_stride_elems) {
    if (engine == nullptr) return fail(CORR_ERR_ARG, "engine is NULL");
    if (d_spectra == nullptr || station_stride_elems == nullptr)
        return fail(CORR_ERR_ARG, "output pointer is NULL");
    *d_spectra = as_engine(engine)->spectra();
    *station_stride_elems = as_engine(engine)->spectra_station_stride();
    return CORR_OK;
}

extern "C" corr_status corrx_set_xtile(void *engine, int32_t channels, int32_t spectra_tile) {
    if (engine == nullptr) return fail(CORR_ERR_ARG, "engine is NULL");
    return as_engine(engine)->set_xtile(channels, spectra_tile);
}

extern "C" corr_status corrx_set_xstep(void *engine, int32_t frames_per_step) {
    if (engine == nullptr) return fail(CORR_ERR_ARG, "engine is NULL");
    return as_engine(engine)->set_xstep(frames_per_step);
}

extern "C" corr_status corrx_get_xstep(void *engine, int32_t *frames_per_step) {
    if (engine == nullptr) return fail(CORR_ERR_ARG, "engine is NULL");
    if (frames_per_step == nullptr) return fail(CORR_ERR_ARG, "output pointer is NULL");
    *frames_per_step = as_engine(engine)->xstep();
    return CORR_OK;
}

extern "C" corr_status corrx_get_xtile(void *engine, int32_t *channels, int32_t *spectra_tile) {
    if (engine == nullptr) return fail(CORR_ERR_ARG, "engine is NULL");
    if (channels == nullptr || spectra_tile == nullptr) return fail(CORR_ERR_ARG, "output pointer is NULL");
    *channels = as_engine(engine)->xtile();
    *spectra_tile = as_engine(engine)->spec_tile();
    return CORR_OK;
}

extern "C" corr_status corrx_batch_create(void *const *engines, int32_t n_engines, void **out_batch) {
    if (out_batch == nullptr) return fail(CORR_ERR_ARG, "out_batch is NULL");
    *out_batch = nullptr;
    if (engines == nullptr || n_engines < 1) return fail(CORR_ERR_ARG, "batch: needs at least one engine");
    std::vector<Engine *> e((size_t)n_engines);
    for (int32_t i = 0; i < n_engines; ++i) e[(size_t)i] = as_engine(engines[i]);
    Batch *b = nullptr;
    const corr_status st = Batch::create(e.data(), n_engines, &b);
    if (st != CORR_OK) return st;
    *out_batch = b;
    return CORR_OK;
}

extern "C" corr_status corrx_batch_destroy(void *batch) {
    delete static_cast<Batch *>(batch);
    return CORR_OK;
}

extern "C" corr_status corrx_batch_run(void *batch, const corrx_batch_item *items, int32_t use_graph,
                                       void *stream) {
    if (batch == nullptr) return fail(CORR_ERR_ARG, "batch is NULL");
    return static_cast<Batch *>(batch)->run(items, use_graph != 0, as_stream(stream));
}
