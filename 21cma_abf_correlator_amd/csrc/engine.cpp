// engine.cpp -- host side of the device engine: tables, plans, stage launches.
#include "engine.hpp"

#ifndef CORR_HD
#define CORR_HD __host__ __device__
#endif
#include "fft_body.hpp"
#include "xengine_body.hpp"
#include "xmfma_body.hpp"

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <vector>

namespace corr {

std::string &last_error() {
    static thread_local std::string s;
    return s;
}

corr_status fail(corr_status st, const std::string &msg) {
    last_error() = msg;
    return st;
}

corr_status failf(corr_status st, const char *fmt, ...) {
    char buf[768];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    last_error() = buf;
    return st;
}

corr_status hip_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return CORR_OK;
    return failf(CORR_ERR_CUDA, "%s: %s: %s", what, hipGetErrorName(e), hipGetErrorString(e));
}

bool fft_size_supported(int n_freq) {
    if (n_freq > 16384) return fft::staged_plan(n_freq).m > 0;
    if (n_freq < 512) return false;
    if ((n_freq & (n_freq - 1)) == 0) return true;
    return n_freq == 6144 || n_freq == 10240 || n_freq == 12288 || n_freq == 14336;
}

int fft_twiddle_elems(int n_freq) {
    int p = n_freq & -n_freq;  // largest power of two dividing n_freq
    if (p > 16384) p = 16384;
    return p == n_freq ? n_freq : n_freq + p;
}

// Frames per batch of the staged transform's two-kernel form: at most 256, and
// fewer where its buffer (8 bytes per retained channel, input and frame) would
// pass 1 GiB (at 24 inputs: 256 frames up to n_freq 20480, 170 at 32768, 5 at
// 2^20). 256-frame batches measured 1-3% faster than 64 and 64 up to 12%
// faster than 16 (fewer launches, each filling the device longer).
constexpr int kStageFrames = 256;
constexpr size_t kStageBytes = size_t(1) << 30;

corr_status Engine::create(const EngineShape &shape, Engine **out) {
    if (out == nullptr) return fail(CORR_ERR_ARG, "engine: out is NULL");
    *out = nullptr;
    if (shape.n_station < 1) return fail(CORR_ERR_ARG, "engine: n_station must be >= 1");
    if (shape.chunk_frames < 1) return fail(CORR_ERR_ARG, "engine: chunk_frames must be >= 1");
    if (!fft_size_supported(shape.n_freq))
        return failf(CORR_ERR_NOT_IMPL,
                     "n_freq=%d: this build transforms n_freq = 2^k in [512, 16384], 6144, 10240, "
                     "12288, 14336, and the multiples of 2048 up to 1048576 (FFT length 2*n_freq)",
                     shape.n_freq);
    if (shape.form != Form::kFP32 && shape.form != Form::kFP16 && shape.form != Form::kINT8)
        return failf(CORR_ERR_ARG, "engine: unknown form %d", (int)shape.form);
    if (shape.form != Form::kFP32 && shape.n_station > kMaxStationMfma)
        return failf(CORR_ERR_NOT_IMPL, "engine: the FP16/int8 forms take at most %d inputs (got %d)",
                     kMaxStationMfma, shape.n_station);
    // the F-engine addresses one input's spectra with 32-bit element offsets
    if ((size_t)shape.chunk_frames * (size_t)shape.n_freq > (size_t)UINT32_MAX)
        return failf(CORR_ERR_ARG, "engine: chunk_frames=%d x n_freq=%d: one input's spectra must stay under 2^32 values",
                     shape.chunk_frames, shape.n_freq);
    Engine *e = new (std::nothrow) Engine();
    if (e == nullptr) return fail(CORR_ERR_NOMEM, "engine: out of host memory");
    e->shape_ = shape;
    corr_status st = e->init();
    if (st != CORR_OK) {
        delete e;
        return st;
    }
    *out = e;
    return CORR_OK;
}

corr_status Engine::init() {
    const int N = shape_.n_station;
    const int nf = shape_.n_freq;
    n_baseline_ = N * (N + 1) / 2;

    corr_status st;
    if ((st = hip_check(hipMalloc(reinterpret_cast<void **>(&d_tw_), sizeof(float2) * (size_t)fft_twiddle_elems(nf)),
                        "hipMalloc(twiddles)")) != CORR_OK ||
        (st = hip_check(hipMalloc(reinterpret_cast<void **>(&d_wl_), sizeof(float2) * nf),
                        "hipMalloc(r2c twiddles)")) != CORR_OK)
        return st;
    // The spectrum buffer also holds cal_frames_ frames of complex64 spectra
    // during calibration.
    const size_t elem = spectra_elem_bytes();
    cal_frames_ = (int)((size_t)shape_.chunk_frames * elem / sizeof(float2));
    if (cal_frames_ < 1) cal_frames_ = 1;
    size_t per_station = (size_t)shape_.chunk_frames * elem;
    if (per_station < (size_t)cal_frames_ * sizeof(float2)) per_station = (size_t)cal_frames_ * sizeof(float2);
    const size_t spec_bytes = (size_t)N * (size_t)nf * per_station;
    if ((st = hip_check(hipMalloc(&d_spec_, spec_bytes), "hipMalloc(spectra)")) != CORR_OK)
        return st;
    if (fft::staged_plan(nf).m > 0) {
        const size_t per_frame = sizeof(float2) * (size_t)N * (size_t)nf;
        const size_t fit = kStageBytes / per_frame < 1 ? 1 : kStageBytes / per_frame;
        stage_frames_ = shape_.chunk_frames < kStageFrames ? shape_.chunk_frames : kStageFrames;
        if ((size_t)stage_frames_ > fit) stage_frames_ = (int)fit;
        if (stage_frames_ > 65535 / N) stage_frames_ = 65535 / N;  // the two-kernel grids' frames x stations
        if (stage_frames_ < 1)
            return failf(CORR_ERR_ARG, "engine: n_station=%d above the staged transform's 65535 at n_freq %d", N, nf);
        const size_t stage = sizeof(float2) * (size_t)N * (size_t)stage_frames_ * (size_t)nf;
        if ((st = hip_check(hipMalloc(reinterpret_cast<void **>(&d_stage_), stage), "hipMalloc(staged transform)")) !=
            CORR_OK)
            return st;
    }
    if ((st = hip_check(hipMalloc(reinterpret_cast<void **>(&d_valid_),
                                  sizeof(int32_t) * (size_t)N * (size_t)shape_.chunk_frames),
                        "hipMalloc(validity)")) != CORR_OK)
        return st;

    // X-engine plan: stations in blocks of 4, lower-triangle block pairs.
    const xeng::Plan plan = xeng::make_plan(N);
    n_bpairs_ = (int)plan.bpairs.size();
    x_npair_ = plan.npair;
    x_waves_ = plan.waves;
    x_groups_ = plan.groups;
    x_frames_lds_ = plan.frames_lds;
    if ((st = hip_check(hipMalloc(reinterpret_cast<void **>(&d_bpairs_), sizeof(int2) * plan.bpairs.size()),
                        "hipMalloc(block pairs)")) != CORR_OK)
        return st;
    if ((st = hip_check(hipMemcpy(d_bpairs_, plan.bpairs.data(), sizeof(int2) * plan.bpairs.size(),
                                  hipMemcpyHostToDevice),
                        "hipMemcpy(block pairs)")) != CORR_OK)
        return st;

    // Baseline table: b -> (j, k), j <= k, j outer (reference corr_cuda.cpp:193-204).
    std::vector<int2> pairs;
    pairs.reserve((size_t)n_baseline_);
    for (int j = 0; j < N; ++j)
        for (int k = j; k < N; ++k) pairs.push_back(make_int2(j, k));
    if ((st = hip_check(hipMalloc(reinterpret_cast<void **>(&d_pairs_), sizeof(int2) * pairs.size()),
                        "hipMalloc(baseline pairs)")) != CORR_OK ||
        (st = hip_check(hipMemcpy(d_pairs_, pairs.data(), sizeof(int2) * pairs.size(), hipMemcpyHostToDevice),
                        "hipMemcpy(baseline pairs)")) != CORR_OK)
        return st;

    if (shape_.form != Form::kFP32) {
        const size_t g = (size_t)N * (size_t)nf;
        if ((st = hip_check(hipMalloc(reinterpret_cast<void **>(&d_gain_), sizeof(float) * g),
                            "hipMalloc(gains)")) != CORR_OK ||
            (st = hip_check(hipMalloc(reinterpret_cast<void **>(&d_power_), sizeof(float) * g),
                            "hipMalloc(power)")) != CORR_OK ||
            (st = hip_check(hipMalloc(reinterpret_cast<void **>(&d_count_), sizeof(int32_t) * (size_t)N),
                            "hipMalloc(counts)")) != CORR_OK)
            return st;
        const float one = 1.0f;
        uint32_t one_bits;
        std::memcpy(&one_bits, &one, 4);
        if ((st = hip_check(launch_fill(d_gain_, g, one_bits, nullptr), "fill_kernel(gains)")) != CORR_OK)
            return st;
    }
    // Every form runs its X-engine on the matrix cores up to 48 inputs (96 Gram
    // rows); the FP32 form beyond that keeps the vector X-engine.
    x_mfma_ = N <= kMaxStationMfma;
    if (!x_mfma_) {
        x_tile_ = 0;  // reported as "no tile": the vector X-engine
        spec_tile_ = nf;
    }
    if (x_mfma_)
        for (int ch : {16, 32})
            for (int kb : {1, 2, 4})
                if (xmfma::tile_fits(N, ch, kb) && nf % ch == 0 &&
                    (st = hip_check(xengine_mfma_prepare((int)shape_.form, N, ch, kb),
                                    "hipFuncSetAttribute(xengine_mfma)")) != CORR_OK)
                    return st;

    if ((st = hip_check(launch_twiddles(d_tw_, d_wl_, nf, nullptr), "twiddle_kernel")) != CORR_OK)
        return st;
    if ((st = hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize(twiddles)")) != CORR_OK) return st;
    return x_mfma_ ? tune_xtile() : CORR_OK;
}

namespace {
struct TileKey {
    int form, n_station, n_freq, frames, device;
    bool operator<(const TileKey &o) const {
        return std::tie(form, n_station, n_freq, frames, device) <
               std::tie(o.form, o.n_station, o.n_freq, o.frames, o.device);
    }
};
/// One X-engine shape: channels per workgroup, spectrum tile, K blocks per step.
struct XShape {
    int ch, tile, kb;
    bool operator==(const XShape &o) const { return ch == o.ch && tile == o.tile && kb == o.kb; }
};
std::mutex g_tile_mu;
std::map<TileKey, XShape> g_tile_cache;
constexpr int kTuneRounds = 7;
constexpr int kTuneFrames = 2048;
}  // namespace

int Engine::frames_per_block() const {
    switch (shape_.form) {
    case Form::kFP16: return xmfma::Traits<1>::kFramesPerBlock;
    case Form::kINT8: return xmfma::Traits<2>::kFramesPerBlock;
    default: return xmfma::Traits<0>::kFramesPerBlock;
    }
}

// The X-engine autotuner. Three choices interact: the channel width of the
// X-engine's workgroups (16, or 32 up to 24 inputs), the channel tile of the
// spectrum order (that width, whose runs the X-engine reads contiguously, or
// n_freq, the plain order the F-engine writes fastest) and the time-K of a
// staged step (1, 2 or 4 MFMA K extents of frames, where the staging fits the
// LDS and a lane stages at most two items: deeper steps halve or quarter the
// barriers per frame and put more bytes in flight per step, at the price of
// LDS and registers). Every combination that fits the shape is timed on the
// engine's own buffers: F-engine over min(chunk_frames, kTuneFrames) frames of
// hashed int8 noise (a transform of zeros draws less power than one of real
// samples, and the part's clock follows its power) and X-engine over them, the
// launch size itself (at fewer frames the
// X-engine's fixed cost per launch hides what the spectrum order changes); one
// warm-up round of each, then kTuneRounds rounds in which every candidate takes
// its turn, timed one by one; a candidate's time is its fastest round, so that
// one disturbed round does not decide. The fastest candidate wins. The
// choice is cached per (form, inputs, n_freq, frames timed, device).
corr_status Engine::tune_xtile() {
    const int N = shape_.n_station, nf = shape_.n_freq;
    std::vector<XShape> cand;
    for (int kb : {1, 2, 4})
        for (int ch : {16, 32})
            if (xmfma::tile_fits(N, ch, kb) && nf % ch == 0) {
                cand.push_back({ch, ch, kb});
                cand.push_back({ch, nf, kb});
            }
    if (cand.empty()) return failf(CORR_ERR_NOT_IMPL, "xengine: no MFMA tile fits %d inputs", N);
    x_tile_ = cand[0].ch;
    spec_tile_ = cand[0].tile;
    x_kb_ = cand[0].kb;
    // CORRX_XTILE=<channels>,<spectrum tile>[,<frames per step>] (tile 0:
    // n_freq, the plain order; frames per step 0 or absent: one K extent) pins
    // the choice wherever it fits the shape, with no timing at creation; where
    // it does not fit, the tuner runs. "0,0" pins the FP32 form's vector
    // X-engine.
    if (const char *env = std::getenv("CORRX_XTILE")) {
        int ch = 0, tl = -1, fps = 0;
        const int got = std::sscanf(env, "%d,%d,%d", &ch, &tl, &fps);
        if (got >= 2) {
            if (tl == 0) tl = nf;
            if (ch == 0 && shape_.form == Form::kFP32) return set_xtile(0, nf);
            const int fb = frames_per_block();
            const int kb = fps <= 0 ? 1 : (fps % fb == 0 ? fps / fb : -1);
            for (const auto &c : cand)
                if (c.ch == ch && c.tile == tl && c.kb == kb) {
                    x_tile_ = c.ch;
                    spec_tile_ = c.tile;
                    x_kb_ = c.kb;
                    return CORR_OK;
                }
        }
    }
    int dev = 0;
    corr_status st;
    if ((st = hip_check(hipGetDevice(&dev), "hipGetDevice")) != CORR_OK) return st;
    const int frames = shape_.chunk_frames < kTuneFrames ? shape_.chunk_frames : kTuneFrames;
    const TileKey key{(int)shape_.form, N, nf, frames, dev};
    {
        std::lock_guard<std::mutex> lk(g_tile_mu);
        auto it = g_tile_cache.find(key);
        if (it != g_tile_cache.end()) {
            x_tile_ = it->second.ch;
            spec_tile_ = it->second.tile;
            x_kb_ = it->second.kb;
            return CORR_OK;
        }
    }
    // One station's frames, read by every station (station pitch 0): the
    // F-engine is VALU-bound, and the candidates differ in how the spectra are
    // written and read, not in how the input is.
    const size_t in_bytes = (size_t)frames * 2u * (size_t)nf;
    const size_t integ_bytes = (size_t)n_baseline_ * (size_t)nf * integrator_elem_bytes();
    void *integ = nullptr, *in = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float best_ms = 0.0f;
    XShape best = cand[0];
    if ((st = hip_check(hipMalloc(&integ, integ_bytes), "hipMalloc(tuning integrator)")) == CORR_OK &&
        (st = hip_check(hipMalloc(&in, in_bytes), "hipMalloc(tuning input)")) == CORR_OK &&
        (st = hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate")) == CORR_OK &&
        (st = hip_check(hipEventCreate(&e0), "hipEventCreate")) == CORR_OK &&
        (st = hip_check(hipEventCreate(&e1), "hipEventCreate")) == CORR_OK &&
        (st = hip_check(launch_noise(in, in_bytes / 4, s), "noise_kernel(input)")) == CORR_OK &&
        (st = hip_check(launch_fill(integ, integ_bytes / 4, 0u, s), "fill_kernel(integrator)")) == CORR_OK) {
        auto round = [&](const XShape &c) {
            corr_status r = transform((int)shape_.form, static_cast<const int8_t *>(in), 0, nullptr, 0, frames,
                                      d_spec_, spectra_station_stride(), c.tile, s);
            if (r != CORR_OK) return r;
            return hip_check(launch_xengine_mfma((int)shape_.form, N, c.ch, c.kb, d_spec_, spectra_station_stride(),
                                                 nf, c.tile, frames, integ, s),
                             "xengine_mfma_kernel");
        };
        // the candidates take turns round by round, so that a drift of the clock
        // over the tuning (the part heats up) reaches all of them alike
        std::vector<float> cand_ms(cand.size(), 0.0f);
        for (size_t i = 0; i < cand.size() && st == CORR_OK; ++i) st = round(cand[i]);  // warm-up
        for (int r = 0; r < kTuneRounds && st == CORR_OK; ++r)
            for (size_t i = 0; i < cand.size() && st == CORR_OK; ++i) {
                float ms = 0.0f;
                if ((st = hip_check(hipEventRecord(e0, s), "hipEventRecord")) != CORR_OK ||
                    (st = round(cand[i])) != CORR_OK ||
                    (st = hip_check(hipEventRecord(e1, s), "hipEventRecord")) != CORR_OK ||
                    (st = hip_check(hipEventSynchronize(e1), "hipEventSynchronize")) != CORR_OK ||
                    (st = hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) != CORR_OK)
                    break;
                if (r == 0 || ms < cand_ms[i]) cand_ms[i] = ms;
            }
        if (st == CORR_OK)
            for (size_t i = 0; i < cand.size(); ++i)
                if (i == 0 || cand_ms[i] < best_ms) {
                    best_ms = cand_ms[i];
                    best = cand[i];
                }
    }
    if (s) hipStreamSynchronize(s);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (s) hipStreamDestroy(s);
    hipFree(integ);
    hipFree(in);
    if (st != CORR_OK) return st;
    x_tile_ = best.ch;
    spec_tile_ = best.tile;
    x_kb_ = best.kb;
    std::lock_guard<std::mutex> lk(g_tile_mu);
    g_tile_cache[key] = best;
    return CORR_OK;
}

corr_status Engine::set_xtile(int channels, int spec_tile) {
    if (shape_.form == Form::kFP32 && channels == 0 && spec_tile == shape_.n_freq) {
        // the FP32 form's vector X-engine, on spectra in the plain order
        x_mfma_ = false;
        x_tile_ = 0;
        spec_tile_ = shape_.n_freq;
        return CORR_OK;
    }
    if (shape_.n_station > kMaxStationMfma)
        return failf(CORR_ERR_ARG, "set_xtile: %d inputs take the vector X-engine, which has no tile",
                     shape_.n_station);
    if (!xmfma::tile_fits(shape_.n_station, channels) || shape_.n_freq % channels != 0)
        return failf(CORR_ERR_ARG, "set_xtile: %d channels per workgroup do not fit %d inputs / n_freq %d",
                     channels, shape_.n_station, shape_.n_freq);
    if (spec_tile != channels && spec_tile != shape_.n_freq)
        return failf(CORR_ERR_ARG, "set_xtile: the spectrum tile must be %d or n_freq (%d), not %d", channels,
                     shape_.n_freq, spec_tile);
    x_mfma_ = true;
    x_tile_ = channels;
    spec_tile_ = spec_tile;
    // a step depth the new width cannot hold falls back to one K extent
    if (!xmfma::tile_fits(shape_.n_station, channels, x_kb_)) x_kb_ = 1;
    return CORR_OK;
}

corr_status Engine::set_xstep(int frames_per_step) {
    if (!x_mfma_) return fail(CORR_ERR_ARG, "set_xstep: the vector X-engine has no step depth");
    const int fb = frames_per_block();
    const int kb = frames_per_step / fb;
    if (frames_per_step <= 0 || kb * fb != frames_per_step || !xmfma::kblocks_valid(kb))
        return failf(CORR_ERR_ARG, "set_xstep: frames per step must be %d, %d or %d in this form, not %d", fb,
                     2 * fb, 4 * fb, frames_per_step);
    if (!xmfma::tile_fits(shape_.n_station, x_tile_, kb))
        return failf(CORR_ERR_ARG, "set_xstep: %d frames per step do not fit %d inputs at %d channels per workgroup",
                     frames_per_step, shape_.n_station, x_tile_);
    x_kb_ = kb;
    return CORR_OK;
}

Engine::~Engine() {
    if (xs_ != nullptr) {
        hipStreamSynchronize(xs_);
        hipStreamDestroy(xs_);
    }
    for (int i = 0; i < 2; ++i) {
        if (ev_f_[i] != nullptr) hipEventDestroy(ev_f_[i]);
        if (ev_x_[i] != nullptr) hipEventDestroy(ev_x_[i]);
    }
    hipFree(d_spec_b_);
    hipFree(d_stage_);
    hipFree(d_tw_);
    hipFree(d_wl_);
    hipFree(d_spec_);
    hipFree(d_valid_);
    hipFree(d_bpairs_);
    hipFree(d_pairs_);
    hipFree(d_gain_);
    hipFree(d_power_);
    hipFree(d_count_);
}

size_t Engine::spectra_elem_bytes() const {
    switch (shape_.form) {
    case Form::kFP16: return 4;
    case Form::kINT8: return 2;
    default: return sizeof(float2);
    }
}

size_t Engine::integrator_elem_bytes() const {
    switch (shape_.form) {
    case Form::kFP16: return 16;
    case Form::kINT8: return 8;
    default: return sizeof(float2);
    }
}

corr_status Engine::fengine(const int8_t *d_in, size_t station_pitch, const uint8_t *d_mask,
                            size_t mask_pitch, int mask_bit0, int frames, hipStream_t s) {
    return fengine_to(d_spec_, d_in, station_pitch, d_mask, mask_pitch, mask_bit0, frames, s);
}

corr_status Engine::xengine(void *integ, int frames, hipStream_t s) {
    return xengine_from(d_spec_, integ, frames, s);
}

corr_status Engine::ensure_pipeline() {
    if (xs_ != nullptr) return CORR_OK;
    const size_t bytes = (size_t)shape_.n_station * spectra_station_stride() * spectra_elem_bytes();
    corr_status st;
    if (d_spec_b_ == nullptr &&
        (st = hip_check(hipMalloc(&d_spec_b_, bytes), "hipMalloc(second spectrum buffer)")) != CORR_OK)
        return st;
    for (int i = 0; i < 2; ++i) {
        if (ev_f_[i] == nullptr &&
            (st = hip_check(hipEventCreateWithFlags(&ev_f_[i], hipEventDisableTiming), "hipEventCreate")) != CORR_OK)
            return st;
        if (ev_x_[i] == nullptr &&
            (st = hip_check(hipEventCreateWithFlags(&ev_x_[i], hipEventDisableTiming), "hipEventCreate")) != CORR_OK)
            return st;
    }
    return hip_check(hipStreamCreateWithFlags(&xs_, hipStreamNonBlocking), "hipStreamCreate(X-engine)");
}

// Chunk k: F-engine into buffer k%2 on `s` (after the X-engine of chunk k-2,
// which read that buffer), then the X-engine on xs_ after it. The X-engines
// run in chunk order on xs_, so the integrator sees one writer at a time.
corr_status Engine::integrate(const int8_t *d_in, size_t station_pitch, const uint8_t *d_mask,
                              size_t mask_pitch, int mask_bit0, int64_t frames, void *integ, hipStream_t s) {
    if (frames < 0) return fail(CORR_ERR_ARG, "integrate: negative frame count");
    if (d_in == nullptr) return fail(CORR_ERR_ARG, "integrate: input is NULL");
    if (integ == nullptr) return fail(CORR_ERR_ARG, "integrate: integrator is NULL");
    if (station_pitch / frame_bytes() < (uint64_t)frames)
        return fail(CORR_ERR_ARG, "integrate: station pitch shorter than the frames it holds");
    if (mask_bit0 < 0) return fail(CORR_ERR_ARG, "integrate: negative mask bit offset");
    if (d_mask != nullptr && mask_pitch * 8u < (uint64_t)mask_bit0 + (uint64_t)frames)
        return fail(CORR_ERR_ARG, "integrate: mask pitch shorter than the bits it must hold");
    if (d_mask != nullptr && (uint64_t)mask_bit0 + (uint64_t)frames > (uint64_t)INT32_MAX)
        return fail(CORR_ERR_ARG, "integrate: mask bits beyond 2^31");
    if (frames == 0) return CORR_OK;
    const int chunk = shape_.chunk_frames;
    corr_status st = CORR_OK;
    if (!overlaps(frames)) {
        // one stream: the FP32 form's X-engine is VALU-bound like the F-engine,
        // and beside it both run slower than in turn; one run has no overlap
        for (int64_t f0 = 0; f0 < frames && st == CORR_OK; f0 += chunk) {
            const int n = frames - f0 < chunk ? (int)(frames - f0) : chunk;
            const int bit0 = d_mask != nullptr ? mask_bit0 + (int)f0 : 0;
            if ((st = fengine_to(d_spec_, d_in + (size_t)f0 * frame_bytes(), station_pitch, d_mask,
                                 mask_pitch, bit0, n, s)) == CORR_OK)
                st = xengine_from(d_spec_, integ, n, s);
        }
        return st;
    }
    if ((st = ensure_pipeline()) != CORR_OK) return st;
    void *const spec[2] = {d_spec_, d_spec_b_};
    int64_t k = 0;
    int last = -1;  // buffer whose X-engine was enqueued last
    for (int64_t f0 = 0; f0 < frames && st == CORR_OK; f0 += chunk, ++k) {
        const int b = (int)(k & 1);
        const int n = frames - f0 < chunk ? (int)(frames - f0) : chunk;
        const int bit0 = d_mask != nullptr ? mask_bit0 + (int)f0 : 0;
        if ((k < 2 || (st = hip_check(hipStreamWaitEvent(s, ev_x_[b], 0), "hipStreamWaitEvent")) == CORR_OK) &&
            (st = fengine_to(spec[b], d_in + (size_t)f0 * frame_bytes(), station_pitch, d_mask,
                             mask_pitch, bit0, n, s)) == CORR_OK &&
            (st = hip_check(hipEventRecord(ev_f_[b], s), "hipEventRecord")) == CORR_OK &&
            (st = hip_check(hipStreamWaitEvent(xs_, ev_f_[b], 0), "hipStreamWaitEvent")) == CORR_OK &&
            (st = xengine_from(spec[b], integ, n, xs_)) == CORR_OK)
            st = hip_check(hipEventRecord(ev_x_[b], xs_), "hipEventRecord");
        if (st == CORR_OK) last = b;
    }
    // Join even after a failed launch, so that `s` never runs ahead of an
    // X-engine already enqueued.
    if (last >= 0) {
        const corr_status j = hip_check(hipStreamWaitEvent(s, ev_x_[last], 0), "hipStreamWaitEvent");
        if (st == CORR_OK) st = j;
    }
    return st;
}

corr_status Engine::fengine_to(void *spec, const int8_t *d_in, size_t station_pitch, const uint8_t *d_mask,
                               size_t mask_pitch, int mask_bit0, int frames, hipStream_t s) {
    if (frames < 0 || frames > shape_.chunk_frames)
        return failf(CORR_ERR_ARG, "fengine: frames=%d outside [0, %d]", frames,
                     shape_.chunk_frames);
    if (d_in == nullptr) return fail(CORR_ERR_ARG, "fengine: input is NULL");
    if (station_pitch < (size_t)frames * frame_bytes())
        return fail(CORR_ERR_ARG, "fengine: station pitch shorter than the frames it holds");
    if (mask_bit0 < 0) return fail(CORR_ERR_ARG, "fengine: negative mask bit offset");
    if (d_mask != nullptr && mask_pitch * 8u < (size_t)(mask_bit0 + frames))
        return fail(CORR_ERR_ARG, "fengine: mask pitch shorter than the bits it must hold");
    // The validity bits are expanded to one int32 per (station, frame) by a
    // separate kernel; the F-engine reads that flag once per workgroup.
    const int32_t *valid = nullptr;
    const size_t valid_stride = (size_t)shape_.chunk_frames;
    if (d_mask != nullptr) {
        corr_status st = hip_check(launch_validity(d_mask, mask_pitch, mask_bit0, frames,
                                                   shape_.n_station, d_valid_, valid_stride, s),
                                   "validity_kernel");
        if (st != CORR_OK) return st;
        valid = d_valid_;
    }
    return transform((int)shape_.form, d_in, station_pitch, valid, valid_stride, frames, spec,
                     spectra_station_stride(), spec_tile_, s);
}

corr_status Engine::transform(int form, const int8_t *d_in, size_t station_pitch, const int32_t *valid,
                              size_t valid_stride, int frames, void *out, size_t out_station_stride, int tile,
                              hipStream_t s) {
    const int N = shape_.n_station, nf = shape_.n_freq;
    if (d_stage_ != nullptr) {
        // the one-kernel form stages nothing: its launches take every frame
        const int batch = fft::staged_one_kernel(form, fft::staged_plan(nf)) ? 65535 : stage_frames_;
        for (int f0 = 0; f0 < frames; f0 += batch) {
            const int n = frames - f0 < batch ? frames - f0 : batch;
            corr_status st = hip_check(launch_fengine_staged(form, nf, d_in, station_pitch, valid, valid_stride, f0, n,
                                                             N, d_tw_, d_wl_, d_gain_, out, out_station_stride, tile,
                                                             d_stage_, s),
                                       "staged transform kernels");
            if (st != CORR_OK) return st;
        }
        return CORR_OK;
    }
    if (form == 0)
        return hip_check(launch_fengine_fp32(nf, d_in, station_pitch, valid, valid_stride, d_tw_, d_wl_,
                                             static_cast<float2 *>(out), out_station_stride, tile, frames, N, s),
                         "fengine_fp32_kernel");
    return hip_check(launch_fengine_form(form, nf, d_in, station_pitch, valid, valid_stride, d_tw_, d_wl_, d_gain_,
                                         out, out_station_stride, tile, frames, N, s),
                     "fengine_form_kernel");
}

corr_status Engine::xengine_from(const void *spec, void *integ, int frames, hipStream_t s) {
    if (integ == nullptr) return fail(CORR_ERR_ARG, "xengine: integrator is NULL");
    if (frames < 0 || frames > shape_.chunk_frames)
        return failf(CORR_ERR_ARG, "xengine: frames=%d outside [0, %d]", frames, shape_.chunk_frames);
    if (!x_mfma_)
        return hip_check(launch_xengine_fp32(static_cast<const float2 *>(spec), spectra_station_stride(),
                                             shape_.n_freq, frames, shape_.n_station, d_bpairs_, n_bpairs_,
                                             x_npair_, x_waves_, x_groups_, x_frames_lds_,
                                             static_cast<float2 *>(integ), s),
                         "xengine_fp32_kernel");
    return hip_check(launch_xengine_mfma((int)shape_.form, shape_.n_station, x_tile_, x_kb_, spec,
                                         spectra_station_stride(), shape_.n_freq, spec_tile_, frames, integ, s),
                     "xengine_mfma_kernel");
}

corr_status Engine::dump(const void *integ, uint64_t frames, float2 *out, hipStream_t s) {
    if (integ == nullptr || out == nullptr) return fail(CORR_ERR_ARG, "dump: NULL buffer");
    if (frames == 0) return fail(CORR_ERR_ARG, "dump: frames must be > 0");
    // The int8 form's int32 integrator holds a frame's largest product sum
    // 2 * 7 * 7 = 98 at most 2^31 / 98 times: longer integrations could have
    // wrapped, so they are refused (the nominal count bounds the frames added).
    if (shape_.form == Form::kINT8 && frames > kInt8MaxFrames)
        return failf(CORR_ERR_ARG, "dump: the int8 form integrates at most %llu frames (int32 sums of products "
                     "up to 98), got %llu", (unsigned long long)kInt8MaxFrames, (unsigned long long)frames);
    return hip_check(launch_dump((int)shape_.form, integ, n_baseline_, shape_.n_freq, d_pairs_, d_gain_,
                                 frames, out, s),
                     "dump_kernel");
}

corr_status Engine::calibrate(const int8_t *d_in, size_t station_pitch, const uint8_t *d_mask,
                              size_t mask_pitch, int mask_bit0, int frames, hipStream_t s) {
    if (shape_.form == Form::kFP32) return fail(CORR_ERR_ARG, "calibrate: the FP32 form has no gains");
    if (frames < 1) return fail(CORR_ERR_ARG, "calibrate: frames must be >= 1");
    if (d_in == nullptr) return fail(CORR_ERR_ARG, "calibrate: input is NULL");
    const int N = shape_.n_station, nf = shape_.n_freq;
    if (station_pitch < (size_t)frames * 2u * (size_t)nf)
        return fail(CORR_ERR_ARG, "calibrate: station pitch shorter than the frames it holds");
    if (mask_bit0 < 0) return fail(CORR_ERR_ARG, "calibrate: negative mask bit offset");
    if (d_mask != nullptr && mask_pitch * 8u < (size_t)mask_bit0 + (size_t)frames)
        return fail(CORR_ERR_ARG, "calibrate: mask pitch shorter than the bits it must hold");
    const size_t g = (size_t)N * (size_t)nf;
    corr_status st;
    if ((st = hip_check(launch_fill(d_power_, g, 0u, s), "fill_kernel(power)")) != CORR_OK ||
        (st = hip_check(launch_fill(d_count_, (size_t)N, 0u, s), "fill_kernel(counts)")) != CORR_OK)
        return st;
    const size_t valid_stride = (size_t)shape_.chunk_frames;
    const size_t cal_stride = (size_t)cal_frames_ * (size_t)nf;
    float2 *spec = static_cast<float2 *>(d_spec_);
    for (int b0 = 0; b0 < frames; b0 += cal_frames_) {
        const int nb = frames - b0 < cal_frames_ ? frames - b0 : cal_frames_;
        const int32_t *valid = nullptr;
        if (d_mask != nullptr) {
            if ((st = hip_check(launch_validity(d_mask, mask_pitch, mask_bit0 + b0, nb, N, d_valid_,
                                                valid_stride, s),
                                "validity_kernel")) != CORR_OK)
                return st;
            valid = d_valid_;
        }
        if ((st = transform(0, d_in + (size_t)b0 * 2u * (size_t)nf, station_pitch, valid, valid_stride, nb, spec,
                            cal_stride, nf, s)) != CORR_OK ||
            (st = hip_check(launch_power(nf, nb, N, spec, cal_stride, valid, valid_stride, d_power_, d_count_, s),
                            "power_kernel")) != CORR_OK)
            return st;
    }
    return hip_check(launch_gains((int)shape_.form, nf, N, d_power_, d_count_, d_gain_, s), "gains_kernel");
}

corr_status Engine::set_gains(const float *d_gains, hipStream_t s) {
    if (shape_.form == Form::kFP32) return fail(CORR_ERR_ARG, "set_gains: the FP32 form has no gains");
    if (d_gains == nullptr) return fail(CORR_ERR_ARG, "set_gains: gains are NULL");
    return hip_check(hipMemcpyAsync(d_gain_, d_gains, sizeof(float) * (size_t)shape_.n_station * shape_.n_freq,
                                    hipMemcpyDeviceToDevice, s),
                     "hipMemcpyAsync(gains)");
}

corr_status Engine::xengine_fp32(float2 *acc, int frames, hipStream_t s) {
    if (shape_.form != Form::kFP32) return fail(CORR_ERR_ARG, "xengine_fp32: the engine is not of the FP32 form");
    if (acc == nullptr) return fail(CORR_ERR_ARG, "xengine: accumulator is NULL");
    return xengine_from(d_spec_, acc, frames, s);
}

corr_status Engine::scale(float2 *acc, size_t n, uint64_t frames, hipStream_t s) {
    if (frames == 0) return fail(CORR_ERR_ARG, "scale: frames must be > 0");
    // 1.0f/(float)frames, then a multiply, in FP32 (reference kernels.cu:890-891).
    const float inv = 1.0f / (float)frames;
    return hip_check(launch_scale(acc, n, inv, s), "scale_kernel");
}

}  // namespace corr
