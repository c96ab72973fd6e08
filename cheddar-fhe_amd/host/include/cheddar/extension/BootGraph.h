// Several bootstraps captured into one HIP graph (config 4's batched form):
// the BootContext's launch sequence for each (input -> output) pair is
// recorded once and replayed with one hipGraphLaunch. Nothing in Boot
// synchronises the host or allocates device memory once the pool holds its
// working set (one eager bootstrap first warms it), and the temporaries it
// frees during capture stay reserved for the graph (DevicePool scope) until
// the graph is destroyed. Inputs and outputs are bound at capture: refill
// the inputs in place and relaunch.
//
// Lanes: with K BootContexts of one parameter set (each with its own
// workspace and prepared transforms), bootstrap i runs on lane i mod K, each
// lane on a stream of its own forked from and joined back to the capture
// stream, so the graph holds K independent chains that the GPU overlaps
// (their launches are latency-bound). Each lane frees into its own pool
// scope, so no temporary passes between lanes; results are identical to
// one lane's. The lanes capture their EvalMod pairs in order (the two-stream
// EvalMod of an eager bootstrap measured slower inside a multi-lane graph:
// 10.4 against 8.8 ms per bootstrap over 4 lanes).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "cheddar/extension/BootContext.h"

namespace cheddar {

template <typename word>
class BootGraph {
  using Ct = Ciphertext<word>;

 public:
  BootGraph(BootContext<word> &ctx, const std::vector<const Ct *> &inputs,
            const std::vector<Ct *> &outputs, const EvkMap<word> &keys, hipStream_t stream)
      : BootGraph(std::vector<BootContext<word> *>{&ctx}, inputs, outputs, keys, stream) {}

  BootGraph(const std::vector<BootContext<word> *> &lanes,
            const std::vector<const Ct *> &inputs, const std::vector<Ct *> &outputs,
            const EvkMap<word> &keys, hipStream_t stream)
      : stream_(stream), scopes_(lanes.size()) {
    if (inputs.size() != outputs.size() || inputs.empty())
      throw Error("BootGraph: one output per input");
    if (!stream) throw Error("BootGraph: capture needs a non-default stream");
    if (lanes.empty()) throw Error("BootGraph: no lane context");
    for (auto *l : lanes)
      if (!l) throw Error("BootGraph: null lane context");
    const size_t k = lanes.size();
    try {
      lane_streams_.push_back(stream);
      for (size_t j = 1; j < k; j++) {
        hipStream_t s = nullptr;
        HipCheck(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        lane_streams_.push_back(s);
        owned_.push_back(s);
      }
      for (size_t j = 0; j < 2 * k; j++) {
        hipEvent_t e = nullptr;
        HipCheck(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreateWithFlags");
        events_.push_back(e);
      }
    } catch (...) {
      Free();
      throw;
    }
    // the graph reuses pooled blocks that eager work on the contexts' own
    // streams may still be using: let all device work finish first
    HipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    std::vector<void *> old(k);
    std::vector<bool> overlap(k);
    for (size_t j = 0; j < k; j++) {
      old[j] = lanes[j]->Stream();
      overlap[j] = lanes[j]->PairOverlap();
      lanes[j]->SetPairOverlap(false);
    }
    auto restore = [&]() {
      DevicePool::Get().EndScope();
      for (size_t j = 0; j < k; j++) {
        lanes[j]->SetStream(old[j]);
        lanes[j]->SetPairOverlap(overlap[j]);
      }
    };
    hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
      try {
        // fork: every lane stream waits for the capture stream's start
        HipCheck(hipEventRecord(events_[0], stream), "hipEventRecord");
        for (size_t j = 1; j < k; j++)
          HipCheck(hipStreamWaitEvent(lane_streams_[j], events_[0], 0), "hipStreamWaitEvent");
        for (size_t j = 0; j < k; j++) {
          DevicePool::Get().BeginScope(&scopes_[j]);
          lanes[j]->SetStream(lane_streams_[j]);
          for (size_t i = j; i < inputs.size(); i += k)
            lanes[j]->Boot(*outputs[i], *inputs[i], keys);
          DevicePool::Get().EndScope();
        }
        // join: the capture stream waits for every lane
        for (size_t j = 1; j < k; j++) {
          HipCheck(hipEventRecord(events_[k + j], lane_streams_[j]), "hipEventRecord");
          HipCheck(hipStreamWaitEvent(stream, events_[k + j], 0), "hipStreamWaitEvent");
        }
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        restore();
        Free();
        throw;
      }
      e = hipStreamEndCapture(stream, &graph_);
    }
    restore();
    if (e != hipSuccess) Free();
    HipCheck(e, "graph capture");
    HipCheck(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
    size_t n = 0;
    HipCheck(hipGraphGetNodes(graph_, nullptr, &n), "hipGraphGetNodes");
    nodes_ = n;
  }

  BootGraph(const BootGraph &) = delete;
  BootGraph &operator=(const BootGraph &) = delete;

  ~BootGraph() {
    (void)hipStreamSynchronize(stream_);
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    Free();
  }

  void Launch() { HipCheck(hipGraphLaunch(exec_, stream_), "hipGraphLaunch"); }
  size_t Nodes() const { return nodes_; }
  size_t Lanes() const { return scopes_.size(); }

 private:
  void Free() {
    for (hipEvent_t e : events_) (void)hipEventDestroy(e);
    events_.clear();
    for (hipStream_t s : owned_) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
    owned_.clear();
    for (auto &sc : scopes_) DevicePool::Get().ReleaseScope(&sc);
  }

  hipStream_t stream_;
  std::vector<hipStream_t> lane_streams_, owned_;
  std::vector<hipEvent_t> events_;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  std::vector<DevicePool::Scope> scopes_;
  size_t nodes_ = 0;
};

}  // namespace cheddar
