// batch.hpp -- one dump of several independent sub-arrays at once.
//
// A Batch holds engines of any shapes and forms (each an independent sub-array:
// its own inputs, n_freq, integration length). One run enqueues, per engine and
// on a branch stream of its own: the integrator zeroing (a fill kernel), the
// whole integration (Engine::integrate: F-engine on the branch, X-engine on the
// engine's stream, two spectrum buffers) and the dump. The branches fork from
// and join back into the caller's stream, so the sub-arrays share the GPU
// concurrently; nothing synchronises the host.
//
// In graph mode the whole run is captured once into a hipGraph and the graph
// is replayed while the run's arguments stay the same (new arguments: a new
// capture). The graph holds kernel launches and stream dependencies only: the
// zeroing is a kernel, not a memset node.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "corrx.h"
#include "engine.hpp"

namespace corr {

class Batch {
  public:
    /// The engines must be distinct and live on the current device; they stay
    /// owned by the caller and must outlive the batch.
    static corr_status create(Engine *const *engines, int n, Batch **out);
    ~Batch();

    int size() const { return (int)eng_.size(); }

    /// One run: items[i] applies to engine i.
    corr_status run(const corrx_batch_item *items, bool graph, hipStream_t s);

  private:
    Batch() = default;
    corr_status check_items(const corrx_batch_item *items) const;
    corr_status enqueue(const corrx_batch_item *items, hipStream_t origin);
    corr_status capture(const corrx_batch_item *items);
    void drop_graph();

    std::vector<Engine *> eng_;
    std::vector<hipStream_t> branch_;
    std::vector<hipEvent_t> join_;
    hipEvent_t fork_ = nullptr;
    hipStream_t cap_ = nullptr;          ///< capture origin
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    std::vector<corrx_batch_item> captured_;  ///< the arguments the graph was made with
    bool launched_ = false;                   ///< the graph has been launched
    hipEvent_t done_ = nullptr;               ///< recorded behind its last launch
};

}  // namespace corr
