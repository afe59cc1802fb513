// HIP graphs of whole steps: the caller brackets the launches of one or more steps
// (every entry point of this library enqueues kernels only, with dt, the time and
// the counters on the device, so a step is capturable) with capture_begin /
// capture_end on a non-default stream, then replays the instantiated graph.
#include <hip/hip_runtime.h>

#include "wbmhd_common.hpp"

extern "C" {

int wb_graph_capture_begin(void* stream) {
  if (!stream) return WB_EINVAL;  // the legacy default stream cannot be captured
  const hipError_t e = hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal);
  return e == hipSuccess ? 0 : -(int)e;
}

int wb_graph_capture_end(void* stream, void** graph) {
  if (!stream || !graph) return WB_EINVAL;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g);
  if (e != hipSuccess) return -(int)e;
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return -(int)e;
  *graph = (void*)exec;
  return 0;
}

int wb_graph_launch(void* graph, int times, void* stream) {
  if (!graph || times < 0) return WB_EINVAL;
  for (int k = 0; k < times; ++k) {
    const hipError_t e = hipGraphLaunch((hipGraphExec_t)graph, (hipStream_t)stream);
    if (e != hipSuccess) return -(int)e;
  }
  return 0;
}

int wb_graph_destroy(void* graph) {
  if (!graph) return 0;
  const hipError_t e = hipGraphExecDestroy((hipGraphExec_t)graph);
  return e == hipSuccess ? 0 : -(int)e;
}

}  // extern "C"
