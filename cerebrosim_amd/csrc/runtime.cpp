// C-ABI runtime helpers: error reporting, streams, graph capture (the counterparts of Warp's ScopedCapture,
// capture_launch and synchronize used at lbm.py:437-444 and validate.py:262).
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "common.hpp"

namespace cs {

static thread_local char g_last_error[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

}  // namespace cs

extern "C" const char* cs_last_error(void) { return cs::g_last_error; }

extern "C" int cs_version(void) { return 1 * 10000 + 0 * 100 + 0; }

extern "C" int cs_stream_create(void** stream) {
    CS_REQUIRE(stream != nullptr, "cs_stream_create: null output");
    hipStream_t s = nullptr;
    CS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = reinterpret_cast<void*>(s);
    return CS_OK;
}

extern "C" int cs_stream_destroy(void* stream) {
    if (stream) CS_HIP(hipStreamDestroy(cs::as_stream(stream)));
    return CS_OK;
}

extern "C" int cs_stream_synchronize(void* stream) {
    CS_HIP(hipStreamSynchronize(cs::as_stream(stream)));
    return CS_OK;
}

extern "C" int cs_device_synchronize(void) {
    CS_HIP(hipDeviceSynchronize());
    return CS_OK;
}

extern "C" int cs_graph_begin(void* stream) {
    CS_REQUIRE(stream != nullptr, "cs_graph_begin: capture needs a created stream");
    CS_HIP(hipStreamBeginCapture(cs::as_stream(stream), hipStreamCaptureModeThreadLocal));
    return CS_OK;
}

extern "C" int cs_graph_end(void* stream, void** graph) {
    CS_REQUIRE(stream != nullptr && graph != nullptr, "cs_graph_end: null argument");
    hipGraph_t g = nullptr;
    CS_HIP(hipStreamEndCapture(cs::as_stream(stream), &g));
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        cs::set_error("hipGraphInstantiate failed: %s", hipGetErrorString(e));
        return CS_E_HIP;
    }
    *graph = reinterpret_cast<void*>(exec);
    return CS_OK;
}

extern "C" int cs_graph_launch(void* graph, void* stream) {
    CS_REQUIRE(graph != nullptr, "cs_graph_launch: null graph");
    CS_HIP(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph), cs::as_stream(stream)));
    return CS_OK;
}

extern "C" int cs_graph_destroy(void* graph) {
    if (graph) CS_HIP(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph)));
    return CS_OK;
}
