// Host-only stand-in for the parts of the HIP runtime API that the layer
// tree and batch engine sources (tree.hip, batch.hip) call, so that their
// planning logic compiles with g++ and runs on the CPU: "device" memory is
// host memory, streams run work immediately, a captured graph records the
// launches issued between begin and end of capture and replays them (a
// stream that waits on an event recorded in a capture joins that capture,
// and leaves it when the capturing stream waits on its event).
#pragma once
#include <cstddef>
#include <cstdint>

typedef enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 } hipError_t;
typedef struct EmuStream *hipStream_t;
typedef struct EmuEvent *hipEvent_t;
typedef struct EmuGraph *hipGraph_t;
typedef struct EmuGraph *hipGraphExec_t;
typedef struct EmuNode *hipGraphNode_t;
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
enum { hipStreamNonBlocking = 1 };
enum hipStreamCaptureMode { hipStreamCaptureModeThreadLocal = 1 };
enum { hipEventDisableTiming = 2 };
struct hipDeviceProp_t {
  char name[256];
  char gcnArchName[256];
  int multiProcessorCount;
  int warpSize;
};

hipError_t hipSetDevice(int);
hipError_t hipGetDevice(int *);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *, int);
const char *hipGetErrorString(hipError_t);
hipError_t hipMalloc(void **, size_t);
template <class T>
hipError_t hipMalloc(T **p, size_t n) { return hipMalloc(reinterpret_cast<void **>(p), n); }
hipError_t hipFree(void *);
hipError_t hipMemset(void *, int, size_t);
hipError_t hipMemsetAsync(void *, int, size_t, hipStream_t);
hipError_t hipMemcpy(void *, const void *, size_t, hipMemcpyKind);
hipError_t hipMemcpyAsync(void *, const void *, size_t, hipMemcpyKind, hipStream_t);
hipError_t hipStreamCreateWithFlags(hipStream_t *, unsigned);
hipError_t hipStreamDestroy(hipStream_t);
hipError_t hipStreamSynchronize(hipStream_t);
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode);
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *);
hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, hipGraphNode_t *, char *, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t);
hipError_t hipGraphExecDestroy(hipGraphExec_t);
hipError_t hipGraphDestroy(hipGraph_t);
hipError_t hipEventCreate(hipEvent_t *);
hipError_t hipEventCreateWithFlags(hipEvent_t *, unsigned);
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned);
hipError_t hipEventDestroy(hipEvent_t);
hipError_t hipEventRecord(hipEvent_t, hipStream_t);
hipError_t hipEventSynchronize(hipEvent_t);
hipError_t hipEventElapsedTime(float *, hipEvent_t, hipEvent_t);

// emulator hooks: run `fn` now, or record it when the stream is capturing
void emu_submit(hipStream_t st, void (*fn)(void *), void *arg, size_t arg_bytes);
// bytes currently allocated through hipMalloc
long long emu_live_bytes();
// hipStreamWaitEvent calls on an event of another stream (forks and joins)
long long emu_cross_stream_waits();
// tiles run by momentum + next continuity launches
long long emu_fused_tiles();
