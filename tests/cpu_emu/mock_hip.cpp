// Implementation of the stand-in HIP runtime (see mock_hip/hip/hip_runtime.h).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>

struct MockStream {
    int id;
};
struct MockEvent {
    int recorded;
};

extern "C" {
hipError_t hipMalloc(void **p, size_t n) {
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipFree(void *p) {
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void *p) { return hipFree(p); }
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
    std::memcpy(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t) {
    return hipMemcpy(d, s, n, k);
}
hipError_t hipMemcpy2DAsync(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h,
                            hipMemcpyKind, hipStream_t) {
    if (w > dp || w > sp) return hipErrorInvalidValue;
    for (size_t r = 0; r < h; ++r)
        std::memcpy(static_cast<char *>(d) + r * dp, static_cast<const char *>(s) + r * sp, w);
    return hipSuccess;
}
hipError_t hipMemset(void *d, int v, size_t n) {
    std::memset(d, v, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t) { return hipMemset(d, v, n); }
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipGetDeviceCount(int *n) {
    *n = 1;
    return hipSuccess;
}
hipError_t hipDeviceSynchronize(void) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
    *s = new MockStream{0};
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t e, unsigned) {
    return e ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) {
    *e = new MockEvent{0};
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
    e->recorded = 1;
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) { return e ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipEventCreate(hipEvent_t *e) { return hipEventCreateWithFlags(e, 0); }
// Every stream operation has completed when it returns: elapsed times are zero.
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
    if (ms == nullptr || a == nullptr || b == nullptr) return hipErrorInvalidValue;
    *ms = 0.0f;
    return hipSuccess;
}
hipError_t hipGetDevice(int *d) {
    if (d == nullptr) return hipErrorInvalidValue;
    *d = 0;
    return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e) { return e ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipLaunchHostFunc(hipStream_t, hipHostFn_t fn, void *user) {
    fn(user);
    return hipSuccess;
}
hipError_t hipGetLastError(void) { return hipSuccess; }
const char *hipGetErrorName(hipError_t) { return "hipErrorMock"; }
const char *hipGetErrorString(hipError_t) { return "mock runtime error"; }
hipError_t hipRuntimeGetVersion(int *v) {
    *v = 0;
    return hipSuccess;
}
hipError_t hipDriverGetVersion(int *v) {
    *v = 0;
    return hipSuccess;
}
// Nothing is deferred here, so no graph can be captured: capture is refused and
// the other graph calls see no graph.
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
    if (g) *g = nullptr;
    return hipErrorNotSupported;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t, hipGraphNode_t *, char *, size_t) {
    if (e) *e = nullptr;
    return hipErrorNotSupported;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { return e ? hipErrorInvalidValue : hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t g) { return g ? hipErrorInvalidValue : hipSuccess; }
}
