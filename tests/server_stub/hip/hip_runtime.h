// tests/server_stub/hip/hip_runtime.h -- a host-only stand-in for the few HIP runtime calls
// drh_server.cpp makes, so its threading logic can be built with g++ under ThreadSanitizer /
// AddressSanitizer on a machine without a GPU (tests/test_server_threads.py).  Test code only:
// it models a stream as a clock of queued work and an event as the time that clock read when
// the event was recorded.
#pragma once
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>

typedef enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorNotReady = 600 } hipError_t;

inline long long stub_now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// A stream: the time its queued work ends.  Work is appended by the stubbed drh_* calls.
struct ihipStream_t {
  std::atomic<long long> busy_until{0};
  void enqueue(long long ns) {
    long long cur = busy_until.load();
    long long next;
    do { next = (cur > stub_now_ns() ? cur : stub_now_ns()) + ns; } while (!busy_until.compare_exchange_weak(cur, next));
  }
};
typedef ihipStream_t* hipStream_t;

struct ihipEvent_t {
  std::atomic<long long> done{0};
  std::atomic<int> magic{0x5EED};
};
typedef ihipEvent_t* hipEvent_t;

inline void stub_check_event(hipEvent_t e, const char* what) {
  if (!e || e->magic.load() != 0x5EED) { fprintf(stderr, "stub: %s on a destroyed event\n", what); abort(); }
}
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = new ihipEvent_t(); return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) {
  stub_check_event(e, "hipEventDestroy");
  e->magic = 0;
  delete e;
  return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  stub_check_event(e, "hipEventRecord");
  e->done = s->busy_until.load();
  return hipSuccess;
}
inline hipError_t hipEventQuery(hipEvent_t e) {
  stub_check_event(e, "hipEventQuery");
  return stub_now_ns() >= e->done.load() ? hipSuccess : hipErrorNotReady;
}
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  stub_check_event(a, "hipEventElapsedTime");
  stub_check_event(b, "hipEventElapsedTime");
  *ms = (float)((b->done.load() - a->done.load()) * 1e-6);
  return hipSuccess;
}
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
#define hipHostMallocDefault 0u
inline hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) {
  *p = std::malloc(bytes);
  return *p ? hipSuccess : hipErrorInvalidValue;
}
inline hipError_t hipHostFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
