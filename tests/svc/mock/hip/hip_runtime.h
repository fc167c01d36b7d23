// A host-only stand-in for the few HIP runtime calls the Poisson service
// makes, for testing the service's scheduling logic on the CPU (test
// infrastructure; never part of the product).  One in-order "device" thread
// executes the stream's commands (copies, kernel launches, event records) a
// little after they are enqueued, so events complete asynchronously.
#pragma once
#include <stddef.h>
#include <stdint.h>

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorNotReady = 600, hipErrorInvalidValue = 1 };
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
struct MockEvent;
struct MockStream;
typedef MockEvent *hipEvent_t;
typedef MockStream *hipStream_t;
typedef void *hipGraph_t;
typedef void *hipGraphExec_t;
#define hipEventDisableTiming 2
#define hipStreamNonBlocking 1
#define hipHostMallocDefault 0
#define __host__
#define __device__
#define __forceinline__ inline

hipError_t hipMalloc(void **p, size_t n);
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind k);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t st);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipStreamCreateWithFlags(hipStream_t *st, unsigned flags);
hipError_t hipStreamSynchronize(hipStream_t st);
hipError_t hipStreamDestroy(hipStream_t st);
hipError_t hipGetDevice(int *d);
hipError_t hipSetDevice(int d);
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t e);
