// Several bootstraps captured into one HIP graph (config 4's batched form):
// the BootContext's launch sequence for each (input -> output) pair is
// recorded once on a capture stream and replayed with one hipGraphLaunch.
// Nothing in Boot synchronises the host or allocates device memory once the
// pool holds its working set (one eager bootstrap first warms it), and the
// temporaries it frees during capture stay reserved for the graph
// (DevicePool scope) until the graph is destroyed. Inputs and outputs are
// bound at capture: refill the inputs in place and relaunch.
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
      : stream_(stream) {
    if (inputs.size() != outputs.size() || inputs.empty())
      throw Error("BootGraph: one output per input");
    if (!stream) throw Error("BootGraph: capture needs a non-default stream");
    // the graph reuses pooled blocks that eager work on the context's own
    // stream may still be using: let all device work finish first
    HipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    void *old = ctx.Stream();
    ctx.SetStream(stream);
    DevicePool::Get().BeginScope(&scope_);
    hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
      try {
        for (size_t i = 0; i < inputs.size(); i++) ctx.Boot(*outputs[i], *inputs[i], keys);
      } catch (...) {
        hipGraph_t g;
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        DevicePool::Get().EndScope();
        ctx.SetStream(old);
        throw;
      }
      e = hipStreamEndCapture(stream, &graph_);
    }
    DevicePool::Get().EndScope();
    ctx.SetStream(old);
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
    DevicePool::Get().ReleaseScope(&scope_);
  }

  void Launch() { HipCheck(hipGraphLaunch(exec_, stream_), "hipGraphLaunch"); }
  size_t Nodes() const { return nodes_; }

 private:
  hipStream_t stream_;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  DevicePool::Scope scope_;
  size_t nodes_ = 0;
};

}  // namespace cheddar
