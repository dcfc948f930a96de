// Host-only stand-in for the HIP runtime API used by the batch engine's host code
// (mass-prime-siever_amd/csrc/mps_batch.hip), so that its scheduler, admission and error paths
// run and are checked on the CPU.  Streams are worker threads that run queued operations in
// order; "device" memory is host memory.  Test infrastructure only: the product links the real
// HIP runtime.
#pragma once
#include <cstddef>
#include <cstdint>

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorLaunchFailure = 719, hipErrorUnknown = 999 };
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice,
                     hipMemcpyDefault };
#define hipStreamNonBlocking 1u
#define hipHostMallocDefault 0u

struct MockStream;
struct MockEvent;
struct MockGraph;
struct MockGraphExec;
typedef MockStream* hipStream_t;
typedef MockEvent* hipEvent_t;
typedef MockGraph* hipGraph_t;
typedef MockGraphExec* hipGraphExec_t;
typedef struct MockNode* hipGraphNode_t;

struct uint2 {
    unsigned int x, y;
};

const char* hipGetErrorString(hipError_t e);
hipError_t hipSetDevice(int d);
hipError_t hipMallocBytes(void** p, size_t n);
hipError_t hipFree(void* p);
hipError_t hipHostMallocBytes(void** p, size_t n, unsigned flags);
hipError_t hipHostFree(void* p);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind k, hipStream_t s);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t* err, char* log, size_t n);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);

template <class T>
hipError_t hipMalloc(T** p, size_t n) {
    return hipMallocBytes(reinterpret_cast<void**>(p), n);
}
template <class T>
hipError_t hipHostMalloc(T** p, size_t n, unsigned flags) {
    return hipHostMallocBytes(reinterpret_cast<void**>(p), n, flags);
}
