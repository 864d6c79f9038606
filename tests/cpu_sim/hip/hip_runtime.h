// Host-only stand-in for <hip/hip_runtime.h>, used to compile the runtime's
// HOST logic (srd_engine.hip) with g++ for CPU tests.  Device code paths are
// never executed in this build: the device-side helpers compile to stubs.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>

#define __device__
#define __host__
#define __global__
#define __shared__
#define __forceinline__ inline
#define __launch_bounds__(x)

typedef int hipError_t;
typedef void* hipStream_t;
typedef void* hipEvent_t;
typedef void* hipMemPool_t;
enum { hipSuccess = 0 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice,
                     hipMemcpyHostToHost };
enum { hipStreamNonBlocking = 1, hipHostMallocDefault = 0 };
enum hipMemPoolAttr { hipMemPoolAttrReleaseThreshold };

inline const char* hipGetErrorString(hipError_t) { return "cpu-sim error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (void*)1; return 0; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
// Everything runs in program order on the host: events and waits do nothing
// but report to sim_stream_event (op 0: a record on s, 1: s waits), which
// only the pool test build defines.
extern "C" __attribute__((weak)) void sim_stream_event(int op, hipStream_t s);
enum { hipEventDisableTiming = 2 };
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = (void*)1; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t s) {
  if (sim_stream_event) sim_stream_event(0, s);
  return hipSuccess;
}
inline hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t, unsigned) {
  if (sim_stream_event) sim_stream_event(1, s);
  return hipSuccess;
}
enum hipStreamCaptureStatus { hipStreamCaptureStatusNone = 0, hipStreamCaptureStatusActive = 1 };
inline hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus* st) {
  *st = hipStreamCaptureStatusNone;
  return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
// Simulated device memory (the pool's host-only tests): hipMalloc fails
// once `capacity` bytes (0 = no limit) are handed out.  Memory is host
// malloc'd and never touched by the pool.
struct SimDevice {
  size_t capacity = 0, used = 0;
  long mallocs = 0, frees = 0, syncs = 0;
  std::map<void*, size_t> blocks;
};
// Defined by the pool test build (pool_sim.cpp): one instance per loaded
// copy of that library.  Other host-only builds never allocate through it.
SimDevice& sim_device();
enum { hipErrorOutOfMemory = 2 };
inline hipError_t hipMalloc(void** p, size_t n) {
  SimDevice& d = sim_device();
  if (d.capacity && d.used + n > d.capacity) return hipErrorOutOfMemory;
  *p = std::malloc(n ? n : 1);
  d.used += n;
  d.blocks[*p] = n;
  d.mallocs++;
  return hipSuccess;
}
inline hipError_t hipFree(void* p) {
  SimDevice& d = sim_device();
  auto it = d.blocks.find(p);
  if (it == d.blocks.end()) std::abort();  // not a block hipMalloc handed out
  d.used -= it->second;
  d.blocks.erase(it);
  d.frees++;
  std::free(p);
  return hipSuccess;
}
inline hipError_t hipDeviceSynchronize() {
  sim_device().syncs++;
  return hipSuccess;
}
inline hipError_t hipMemGetInfo(size_t* fr, size_t* tot) {
  *fr = *tot = size_t(1) << 34;
  return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  if (n) std::memmove(d, s, n);
  return hipSuccess;
}

// Graph capture: the host-only build never captures (the tests run the
// device-resident loop with SRDATALOG_DEVLOOP_GRAPH=0); these only compile.
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
typedef void* hipGraphNode_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal,
                            hipStreamCaptureModeRelaxed };
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { std::abort(); }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { std::abort(); }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, hipGraphNode_t*, char*, size_t) {
  std::abort();
}
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { std::abort(); }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }

struct srd_dim3 { unsigned x = 0, y = 0, z = 0; };
static srd_dim3 threadIdx, blockIdx, blockDim, gridDim;

inline unsigned __builtin_amdgcn_mbcnt_hi(unsigned, unsigned) { std::abort(); }
inline unsigned __builtin_amdgcn_mbcnt_lo(unsigned, unsigned) { std::abort(); }
inline unsigned long long __ballot(int) { std::abort(); }
template <class T> inline T __shfl_up(T, int, int) { std::abort(); }
template <class T> inline T __shfl_xor(T, int, int) { std::abort(); }
template <class T> inline T __shfl(T, int, int) { std::abort(); }
inline void __syncthreads() { std::abort(); }
template <class T> inline T atomicAdd(T*, T) { std::abort(); }
inline unsigned __builtin_amdgcn_readlane(unsigned, int) { std::abort(); }
