// libwfdtd runtime entry points: version, last error, streams and HIP graph
// capture/replay (the native replacement for wp.ScopedCapture and
// wp.capture_launch, warp_fdtd/solver.py:239-258).
#include <cstdarg>

#include "common.hpp"

namespace wfd {

namespace {
thread_local char g_error[512] = "";
}

void set_error(const char* fmt, ...) {
  va_list args;
  va_start(args, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, args);
  va_end(args);
}

}  // namespace wfd

using namespace wfd;

extern "C" {

int wfd_version(void) { return 1; }

const char* wfd_last_error(void) { return g_error; }

int wfd_device_synchronize(void) {
  WFD_HIP(hipDeviceSynchronize());
  return WFD_OK;
}

int wfd_stream_synchronize(void* stream) {
  WFD_HIP(hipStreamSynchronize(as_stream(stream)));
  return WFD_OK;
}

int wfd_stream_create(void** stream) {
  WFD_REQUIRE(stream, "stream_create: null output");
  hipStream_t s;
  WFD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return WFD_OK;
}

int wfd_stream_destroy(void* stream) {
  WFD_REQUIRE(stream, "stream_destroy: null stream");
  WFD_HIP(hipStreamDestroy(as_stream(stream)));
  return WFD_OK;
}

int wfd_graph_begin(void* stream) {
  WFD_REQUIRE(stream, "graph_begin: capture needs a created (non-null) stream");
  WFD_HIP(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal));
  return WFD_OK;
}

int wfd_graph_end(void* stream, void** graph_exec) {
  WFD_REQUIRE(stream && graph_exec, "graph_end: null argument");
  hipGraph_t graph = nullptr;
  WFD_HIP(hipStreamEndCapture(as_stream(stream), &graph));
  hipGraphExec_t exec = nullptr;
  hipError_t err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (err != hipSuccess) {
    set_error("hipGraphInstantiate: %s", hipGetErrorString(err));
    return WFD_ERR_HIP;
  }
  *graph_exec = exec;
  return WFD_OK;
}

int wfd_graph_launch(void* graph_exec, void* stream) {
  WFD_REQUIRE(graph_exec, "graph_launch: null graph");
  WFD_HIP(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph_exec), as_stream(stream)));
  return WFD_OK;
}

int wfd_graph_destroy(void* graph_exec) {
  WFD_REQUIRE(graph_exec, "graph_destroy: null graph");
  WFD_HIP(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph_exec)));
  return WFD_OK;
}

}  // extern "C"
