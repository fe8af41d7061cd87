// A stand-in for the HIP runtime header, for the CPU build of the library used
// only by tests: device memory is host memory, every stream operation runs at
// once in call order, kernels are the emulated kernel bodies. Under
// AddressSanitizer this checks every copy, pointer and launch argument of the
// host code against the exact sizes it allocated.
#pragma once
#include <cstddef>
#include <cstdint>

#include "emu_types.hpp"

#define __host__
#define __device__
#define __global__

typedef int hipError_t;
enum {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorNotReady = 600,
    hipErrorIllegalAddress = 700,
    hipErrorNotSupported = 801,
};
typedef struct MockStream *hipStream_t;
typedef struct MockEvent *hipEvent_t;
typedef void (*hipHostFn_t)(void *);
// Graphs: stream operations run at once here, so there is nothing to capture;
// hipStreamBeginCapture reports hipErrorNotSupported.
typedef struct MockGraph *hipGraph_t;
typedef struct MockGraphExec *hipGraphExec_t;
typedef struct MockGraphNode *hipGraphNode_t;
enum hipStreamCaptureMode {
    hipStreamCaptureModeGlobal = 0,
    hipStreamCaptureModeThreadLocal,
    hipStreamCaptureModeRelaxed,
};
enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
};
#define hipStreamNonBlocking 1u
#define hipEventDisableTiming 2u
#define hipHostMallocDefault 0u

extern "C" {
hipError_t hipMalloc(void **p, size_t n);
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipMemcpy2DAsync(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h,
                            hipMemcpyKind k, hipStream_t st);
hipError_t hipMemset(void *d, int v, size_t n);
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st);
hipError_t hipSetDevice(int d);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipDeviceSynchronize(void);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned f);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned f);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned f);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipGetDevice(int *d);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipLaunchHostFunc(hipStream_t s, hipHostFn_t fn, void *user);
hipError_t hipGetLastError(void);
const char *hipGetErrorName(hipError_t e);
const char *hipGetErrorString(hipError_t e);
hipError_t hipRuntimeGetVersion(int *v);
hipError_t hipDriverGetVersion(int *v);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *err, char *log, size_t n);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);
}
