// Replay of fixed launch sequences as HIP graphs.
//
// The solves of the interior-point step launch the same dozens of small,
// dependent kernels over and over (a triangular solve pair is 32 GEMV
// launches; a preconditioned-CG iteration adds ten more), and issued one by
// one from the host they leave the GPU idle between launches. A sequence is
// captured from its stream once per key (every buffer it reads or writes
// that can change between calls: the graph holds their addresses) and later
// replayed with one graph launch. The sequences hold kernels and
// device-to-device copies only (no memset nodes, no host synchronisation).
// Replay is opt-in (SOVEREIGN_GRAPHS=1): on MI355X it measured no faster
// than launching the same kernels from the host (0.39 ms per FP64 solve pair
// either way), and a run under rocprofv3's kernel tracing once crashed
// inside the runtime's graph path (cause not shown; graphs destroyed while
// a replay could still be in flight were one defect found by reading, and
// clear() now waits for their streams). It is also off while per-kernel event
// timing is on (timing level 1), so that the event pairs see the individual
// launches.
#pragma once

#include "common.hpp"

#include <cstdlib>
#include <string>
#include <vector>

namespace sov {

bool kernel_timing();

inline bool graphs_enabled() {
  const char* e = std::getenv("SOVEREIGN_GRAPHS");
  return e && e[0] == '1' && !kernel_timing();
}

// True while s is being captured (work enqueued now becomes part of a graph).
inline bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &st) == hipSuccess && st == hipStreamCaptureStatusActive;
}

class LaunchGraphs {
 public:
  LaunchGraphs() = default;
  LaunchGraphs(const LaunchGraphs&) = delete;
  LaunchGraphs& operator=(const LaunchGraphs&) = delete;
  ~LaunchGraphs() { clear(); }

  // Destroys the cached graphs, after the streams they were last launched
  // on have finished them (a replay may still be in flight when a caller
  // reallocates the buffers the graphs hold, or tears the owner down).
  void clear() {
    for (Entry& e : entries_) {
      if (e.stream) (void)hipStreamSynchronize(e.stream);
      (void)hipGraphExecDestroy(e.exec);
    }
    entries_.clear();
  }

  // Runs what enqueue() launches on s: eagerly when replay is off or s is
  // already being captured (the launches then join the outer graph),
  // otherwise as the graph captured for (k0, k1, k2, k3), capturing it on
  // first use. enqueue returns a status (SOV_OK on success).
  template <typename F>
  int run(const void* k0, const void* k1, const void* k2, long long k3, hipStream_t s, F&& enqueue) {
    if (!graphs_enabled() || stream_capturing(s)) return enqueue();
    for (Entry& e : entries_)
      if (e.k0 == k0 && e.k1 == k1 && e.k2 == k2 && e.k3 == k3) {
        e.stream = s;
        return hipGraphLaunch(e.exec, s) == hipSuccess ? SOV_OK : fail();
      }
    if (entries_.size() >= kMaxEntries) clear();
    SOV_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &g);
    if (rc != SOV_OK || ec != hipSuccess || !g) {
      if (g) (void)hipGraphDestroy(g);
      if (rc != SOV_OK) return rc;
      set_error(std::string("graph capture failed: ") + hipGetErrorString(ec));
      return SOV_ERR_DEVICE;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiateWithFlags(&exec, g, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess || !exec) {
      set_error(std::string("graph instantiation failed: ") + hipGetErrorString(ei));
      return SOV_ERR_DEVICE;
    }
    entries_.push_back(Entry{k0, k1, k2, k3, exec, s});
    return hipGraphLaunch(exec, s) == hipSuccess ? SOV_OK : fail();
  }

 private:
  static constexpr std::size_t kMaxEntries = 8;
  struct Entry {
    const void* k0;
    const void* k1;
    const void* k2;
    long long k3;
    hipGraphExec_t exec;
    hipStream_t stream;  // where it was last launched
  };
  static int fail() {
    set_error("graph launch failed");
    return SOV_ERR_DEVICE;
  }
  std::vector<Entry> entries_;
};

}  // namespace sov
