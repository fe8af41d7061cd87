// batch.cpp -- concurrent sub-arrays: branch streams, and the hipGraph per dump.
#include "batch.hpp"

#include <cstring>
#include <new>

namespace corr {

corr_status Batch::create(Engine *const *engines, int n, Batch **out) {
    if (out == nullptr) return fail(CORR_ERR_ARG, "batch: out is NULL");
    *out = nullptr;
    if (engines == nullptr || n < 1) return fail(CORR_ERR_ARG, "batch: needs at least one engine");
    for (int i = 0; i < n; ++i) {
        if (engines[i] == nullptr) return failf(CORR_ERR_ARG, "batch: engine %d is NULL", i);
        for (int j = 0; j < i; ++j)
            if (engines[j] == engines[i])
                return failf(CORR_ERR_ARG, "batch: engine %d appears twice (%d and %d)", i, j, i);
    }
    Batch *b = new (std::nothrow) Batch();
    if (b == nullptr) return fail(CORR_ERR_NOMEM, "batch: out of host memory");
    corr_status st = CORR_OK;
    b->eng_.assign(engines, engines + n);
    b->branch_.assign((size_t)n, nullptr);
    b->join_.assign((size_t)n, nullptr);
    for (int i = 0; i < n && st == CORR_OK; ++i)
        if ((st = engines[i]->prepare_integrate()) == CORR_OK &&
            (st = hip_check(hipStreamCreateWithFlags(&b->branch_[(size_t)i], hipStreamNonBlocking),
                            "hipStreamCreate(batch branch)")) == CORR_OK)
            st = hip_check(hipEventCreateWithFlags(&b->join_[(size_t)i], hipEventDisableTiming), "hipEventCreate");
    if (st == CORR_OK)
        st = hip_check(hipEventCreateWithFlags(&b->fork_, hipEventDisableTiming), "hipEventCreate");
    if (st == CORR_OK)
        st = hip_check(hipEventCreateWithFlags(&b->done_, hipEventDisableTiming), "hipEventCreate");
    if (st == CORR_OK)
        st = hip_check(hipStreamCreateWithFlags(&b->cap_, hipStreamNonBlocking), "hipStreamCreate(batch capture)");
    if (st != CORR_OK) {
        delete b;
        return st;
    }
    *out = b;
    return CORR_OK;
}

Batch::~Batch() {
    for (hipStream_t s : branch_)
        if (s != nullptr) hipStreamSynchronize(s);
    if (cap_ != nullptr) hipStreamSynchronize(cap_);
    drop_graph();
    for (hipEvent_t e : join_)
        if (e != nullptr) hipEventDestroy(e);
    for (hipStream_t s : branch_)
        if (s != nullptr) hipStreamDestroy(s);
    if (fork_ != nullptr) hipEventDestroy(fork_);
    if (done_ != nullptr) hipEventDestroy(done_);
    if (cap_ != nullptr) hipStreamDestroy(cap_);
}

// The last launch of the graph may still run on the caller's stream: it ends
// (done_, recorded behind it) before the graph is destroyed.
void Batch::drop_graph() {
    if (exec_ != nullptr && launched_) hipEventSynchronize(done_);
    launched_ = false;
    if (exec_ != nullptr) hipGraphExecDestroy(exec_);
    if (graph_ != nullptr) hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
    captured_.clear();
}

// Everything Engine::integrate and Engine::dump would reject, checked before
// anything is enqueued, so that a capture never stops half-way.
corr_status Batch::check_items(const corrx_batch_item *items) const {
    if (items == nullptr) return fail(CORR_ERR_ARG, "batch: items are NULL");
    for (size_t i = 0; i < eng_.size(); ++i) {
        const corrx_batch_item &it = items[i];
        const EngineShape &sh = eng_[i]->shape();
        if ((it.flags & ~(CORRX_BATCH_ZERO | CORRX_BATCH_DUMP)) != 0)
            return failf(CORR_ERR_ARG, "batch item %zu: unknown flags 0x%x", i, (unsigned)it.flags);
        if (it.d_acc == nullptr) return failf(CORR_ERR_ARG, "batch item %zu: integrator is NULL", i);
        if (it.frames < 0) return failf(CORR_ERR_ARG, "batch item %zu: negative frame count", i);
        if (it.frames > 0) {
            if (it.d_in == nullptr) return failf(CORR_ERR_ARG, "batch item %zu: input is NULL", i);
            if (it.station_pitch / (2u * (size_t)sh.n_freq) < (uint64_t)it.frames)
                return failf(CORR_ERR_ARG, "batch item %zu: station pitch shorter than the frames it holds", i);
            if (it.mask_bit0 < 0) return failf(CORR_ERR_ARG, "batch item %zu: negative mask bit offset", i);
            if (it.d_mask != nullptr && (it.mask_pitch * 8u < (uint64_t)it.mask_bit0 + (uint64_t)it.frames ||
                                         (uint64_t)it.mask_bit0 + (uint64_t)it.frames > (uint64_t)INT32_MAX))
                return failf(CORR_ERR_ARG, "batch item %zu: mask pitch shorter than the bits it must hold", i);
        }
        if ((it.flags & CORRX_BATCH_DUMP) != 0 && (it.d_out == nullptr || it.dump_frames == 0))
            return failf(CORR_ERR_ARG, "batch item %zu: a dump needs an output and dump_frames > 0", i);
    }
    return CORR_OK;
}

corr_status Batch::enqueue(const corrx_batch_item *items, hipStream_t origin) {
    corr_status st = hip_check(hipEventRecord(fork_, origin), "hipEventRecord(fork)");
    size_t forked = 0;
    for (size_t i = 0; i < eng_.size() && st == CORR_OK; ++i) {
        const corrx_batch_item &it = items[i];
        Engine *e = eng_[i];
        hipStream_t b = branch_[i];
        if ((st = hip_check(hipStreamWaitEvent(b, fork_, 0), "hipStreamWaitEvent(fork)")) != CORR_OK) break;
        ++forked;
        const size_t words = (size_t)e->n_baseline() * (size_t)e->shape().n_freq * e->integrator_elem_bytes() / 4u;
        if ((it.flags & CORRX_BATCH_ZERO) != 0)
            st = hip_check(launch_fill(it.d_acc, words, 0u, b), "fill_kernel(integrator)");
        if (st == CORR_OK && it.frames > 0)
            st = e->integrate(it.d_in, it.station_pitch, it.d_mask, it.mask_pitch, it.mask_bit0, it.frames, it.d_acc,
                              b);
        if (st == CORR_OK && (it.flags & CORRX_BATCH_DUMP) != 0)
            st = e->dump(it.d_acc, it.dump_frames, static_cast<float2 *>(it.d_out), b);
    }
    // Join every branch that was forked, also after a failure, so that the
    // origin never runs ahead of work already enqueued on a branch.
    for (size_t i = 0; i < forked; ++i) {
        corr_status j = hip_check(hipEventRecord(join_[i], branch_[i]), "hipEventRecord(join)");
        if (j == CORR_OK) j = hip_check(hipStreamWaitEvent(origin, join_[i], 0), "hipStreamWaitEvent(join)");
        if (st == CORR_OK) st = j;
    }
    return st;
}

corr_status Batch::capture(const corrx_batch_item *items) {
    drop_graph();
    corr_status st = hip_check(hipStreamBeginCapture(cap_, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    if (st != CORR_OK) return st;
    st = enqueue(items, cap_);
    hipGraph_t g = nullptr;
    const corr_status e = hip_check(hipStreamEndCapture(cap_, &g), "hipStreamEndCapture");
    if (st == CORR_OK) st = e;
    if (st == CORR_OK) st = hip_check(hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (st != CORR_OK) {
        if (g != nullptr) hipGraphDestroy(g);
        exec_ = nullptr;
        return st;
    }
    graph_ = g;
    captured_.assign(items, items + eng_.size());
    return CORR_OK;
}

corr_status Batch::run(const corrx_batch_item *items, bool graph, hipStream_t s) {
    corr_status st = check_items(items);
    if (st != CORR_OK) return st;
    if (!graph) return enqueue(items, s);
    const bool same = exec_ != nullptr && captured_.size() == eng_.size() &&
                      std::memcmp(captured_.data(), items, sizeof(corrx_batch_item) * eng_.size()) == 0;
    if (!same && (st = capture(items)) != CORR_OK) return st;
    if ((st = hip_check(hipGraphLaunch(exec_, s), "hipGraphLaunch")) != CORR_OK) return st;
    launched_ = true;
    return hip_check(hipEventRecord(done_, s), "hipEventRecord(graph launch)");
}

}  // namespace corr
