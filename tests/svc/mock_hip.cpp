// Host-only stand-in for the HIP runtime calls of the Poisson service, and
// for the kernels it launches (test infrastructure).  See mock/hip/hip_runtime.h.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>

struct MockEvent {
  bool done = true;
};
struct MockStream {
  int id = 0;
};

namespace {
std::mutex g_mu;
std::condition_variable g_cv, g_done_cv;
std::deque<std::function<void()>> g_queue;
bool g_stop = false;
std::thread *g_dev = nullptr;
long g_enqueued = 0, g_executed = 0;
int g_delay_us = 30;

void device_loop() {
  std::unique_lock<std::mutex> lk(g_mu);
  while (true) {
    g_cv.wait(lk, [] { return g_stop || !g_queue.empty(); });
    if (g_queue.empty() && g_stop) break;
    std::function<void()> f = std::move(g_queue.front());
    g_queue.pop_front();
    lk.unlock();
    if (g_delay_us > 0) std::this_thread::sleep_for(std::chrono::microseconds(g_delay_us));
    f();
    lk.lock();
    g_executed++;
    g_done_cv.notify_all();
  }
}

void ensure_device() {
  if (!g_dev) g_dev = new std::thread(device_loop);
}
}  // namespace

void mock_enqueue(std::function<void()> f) {
  std::lock_guard<std::mutex> g(g_mu);
  ensure_device();
  g_queue.push_back(std::move(f));
  g_enqueued++;
  g_cv.notify_all();
}

// Device time per command, microseconds (0: the device keeps up with the host).
void mock_set_delay(int us) {
  std::lock_guard<std::mutex> g(g_mu);
  g_delay_us = us;
}

// Stops the device thread (all commands executed first).
void mock_shutdown() {
  {
    std::lock_guard<std::mutex> g(g_mu);
    g_stop = true;
    g_cv.notify_all();
  }
  if (g_dev) {
    g_dev->join();
    delete g_dev;
    g_dev = nullptr;
  }
}

void mock_drain() {
  std::unique_lock<std::mutex> lk(g_mu);
  g_done_cv.wait(lk, [] { return g_executed == g_enqueued; });
}

hipError_t hipMalloc(void **p, size_t n) {
  *p = malloc(n ? n : 1);
  return *p ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipFree(void *p) {
  mock_drain();  // frees wait for the device, as in HIP
  free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void *p) { return hipFree(p); }
hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind) {
  mock_drain();
  memcpy(dst, src, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind, hipStream_t) {
  mock_enqueue([=] { memcpy(dst, src, n); });
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) {
  *e = new MockEvent();
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  mock_drain();
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
  {
    std::lock_guard<std::mutex> g(g_mu);
    e->done = false;
  }
  mock_enqueue([e] {
    std::lock_guard<std::mutex> g(g_mu);
    e->done = true;
  });
  return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e) {
  std::lock_guard<std::mutex> g(g_mu);
  return e->done ? hipSuccess : hipErrorNotReady;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
  std::unique_lock<std::mutex> lk(g_mu);
  g_done_cv.wait(lk, [e] { return e->done; });
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *st, unsigned) {
  *st = new MockStream();
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) {
  mock_drain();
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t st) {
  mock_drain();
  delete st;
  return hipSuccess;
}
hipError_t hipGetDevice(int *d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "mock error"; }
