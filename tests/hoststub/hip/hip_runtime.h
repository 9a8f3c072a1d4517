// CPU stand-in for the HIP runtime, for testing the device library's host code without a GPU (test
// infrastructure; never part of the product).  Device memory is host heap, so AddressSanitizer watches it; streams
// run synchronously; graph capture records the launches and a graph launch replays them; kernels are the host
// emulations of emu_kernels.hpp.  A copy, memset or synchronisation issued on a stream that is being captured
// aborts, as those calls must stay outside a capture.
#pragma once
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (&(x))

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyHostToHost, hipMemcpyDefault };
enum { hipStreamNonBlocking = 1, hipEventDisableTiming = 2, hipStreamCaptureModeThreadLocal = 1 };
struct uint4 { unsigned x, y, z, w; };
struct hipDeviceProp_t { char name[256]; char gcnArchName[256]; size_t totalGlobalMem; int multiProcessorCount; };

namespace hipstub {
using Ops = std::vector<std::function<void()>>;
struct Stream { Ops* cap = nullptr; };
struct Graph { Ops ops; };
struct Event { double t = 0; };
inline double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
// device bytes the stand-in reports free (the request server's default HBM budget is 80 % of it)
inline size_t free_bytes() {
  const char* e = std::getenv("HIPSTUB_FREE_MIB");
  return (e ? size_t(std::strtoull(e, nullptr, 10)) : size_t(4096)) << 20;
}
template <typename F, typename... P> void launch(Stream* s, F f, P... p) {
  if (s && s->cap) s->cap->push_back([=]() { f(p...); });
  else f(p...);
}
inline void not_captured(Stream* s) { if (s && s->cap) { std::fprintf(stderr, "hipstub: call inside a capture\n"); std::abort(); } }
}  // namespace hipstub

typedef hipstub::Stream* hipStream_t;
typedef hipstub::Event* hipEvent_t;
typedef hipstub::Graph* hipGraph_t;
typedef hipstub::Graph* hipGraphExec_t;

#define hipLaunchKernelGGL(K, G, B, S, STR, ...) ::hipstub::launch(STR, K, ##__VA_ARGS__)

inline const char* hipGetErrorString(hipError_t) { return "hipstub error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* c) { *c = 1; return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
  std::memset(p, 0, sizeof(*p));
  std::snprintf(p->name, sizeof(p->name), "hipstub");
  std::snprintf(p->gcnArchName, sizeof(p->gcnArchName), "host");
  p->totalGlobalMem = hipstub::free_bytes();
  p->multiProcessorCount = 1;
  return hipSuccess;
}
inline hipError_t hipDeviceGetName(char* n, int cap, int) { std::snprintf(n, cap, "hipstub"); return hipSuccess; }
inline hipError_t hipMemGetInfo(size_t* fr, size_t* tot) { *fr = hipstub::free_bytes(); *tot = *fr; return hipSuccess; }
template <typename T> hipError_t hipMalloc(T** p, size_t n) {
  *p = static_cast<T*>(std::malloc(n ? n : 1));
  return *p ? hipSuccess : hipErrorInvalidValue;
}
inline hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { if (n) std::memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st) {
  hipstub::not_captured(st);
  return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemset(void* d, int v, size_t n) { std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st) { hipstub::not_captured(st); return hipMemset(d, v, n); }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = new hipstub::Stream(); return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t s) { hipstub::not_captured(s); return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = new hipstub::Event(); return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = new hipstub::Event(); return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t = nullptr) { e->t = hipstub::now_ms(); return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) { *ms = float(b->t - a->t); return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t s, int) {
  if (!s || s->cap) { std::fprintf(stderr, "hipstub: capture on a null or capturing stream\n"); std::abort(); }
  s->cap = new hipstub::Ops();
  return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  *g = new hipstub::Graph();
  (*g)->ops = std::move(*s->cap);
  delete s->cap;
  s->cap = nullptr;
  return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t) { *e = new hipstub::Graph(*g); return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
inline hipError_t hipGraphUpload(hipGraphExec_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
  hipstub::not_captured(s);
  for (auto& f : e->ops) f();
  return hipSuccess;
}
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return hipSuccess; }
