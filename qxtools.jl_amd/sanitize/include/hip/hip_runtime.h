// Host-only stand-in for the HIP runtime API used by libqxhip's host code
// (runtime.cpp), for the sanitizer build: a simulated device whose memory is host
// memory, whose streams run nothing and whose kernel launches are checked, not
// executed (sim_device.cpp).  Only the names and signatures the host code uses.
#pragma once
#include <cstddef>
#include <cstdint>

enum hipError_t { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorInvalidDevice = 101 };
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
                     hipMemcpyDeviceToDevice = 3 };
enum hipStreamCaptureStatus { hipStreamCaptureStatusNone = 0, hipStreamCaptureStatusActive = 1 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 };

struct SimStream;
struct SimEvent;
struct SimGraph;
struct SimGraphExec;
typedef SimStream *hipStream_t;
typedef SimEvent *hipEvent_t;
typedef SimGraph *hipGraph_t;
typedef SimGraphExec *hipGraphExec_t;
typedef void *hipGraphNode_t;

#define hipStreamNonBlocking 1u

hipError_t hipSetDevice(int device);
hipError_t hipGetDevice(int *device);
hipError_t hipGetDeviceCount(int *count);
const char *hipGetErrorString(hipError_t e);
hipError_t hipMalloc(void **p, size_t bytes);
hipError_t hipFree(void *p);
hipError_t hipMemGetInfo(size_t *free_bytes, size_t *total_bytes);
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamIsCapturing(hipStream_t s, hipStreamCaptureStatus *st);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *ex, hipGraph_t g, hipGraphNode_t *err_node, char *log, size_t cap);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t ex);
hipError_t hipGraphLaunch(hipGraphExec_t ex, hipStream_t s);
