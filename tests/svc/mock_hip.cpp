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
#include <random>
#include <set>
#include <thread>

struct MockEvent {
  bool done = true;
};

// One in-order queue per stream, executed by its own thread: streams run
// concurrently, as the lanes of the service do on the GPU.
struct MockStream {
  std::deque<std::function<void()>> queue;
  std::thread th;
  long enqueued = 0, executed = 0;
  bool stop = false;
};

namespace {
std::mutex g_mu;
std::condition_variable g_cv, g_done_cv;
std::set<MockStream *> g_streams;
MockStream *g_null = nullptr;  // commands enqueued without a stream
int g_delay_us = 30;
unsigned g_seed = 1;

void stream_loop(MockStream *s, unsigned seed) {
  std::mt19937 rng(seed);
  std::unique_lock<std::mutex> lk(g_mu);
  for (;;) {
    g_cv.wait(lk, [s] { return s->stop || !s->queue.empty(); });
    if (s->queue.empty() && s->stop) break;
    std::function<void()> f = std::move(s->queue.front());
    s->queue.pop_front();
    const int d = g_delay_us;
    lk.unlock();
    // device time per command: the set delay, jittered so that the streams'
    // commands interleave differently from run to run
    if (d > 0) std::this_thread::sleep_for(std::chrono::microseconds(d / 2 + (int)(rng() % (unsigned)(d + 1))));
    f();
    lk.lock();
    s->executed++;
    g_done_cv.notify_all();
  }
}

MockStream *new_stream() {
  MockStream *s = new MockStream();
  g_streams.insert(s);
  s->th = std::thread(stream_loop, s, g_seed++);
  return s;
}

bool all_idle() {
  for (MockStream *s : g_streams)
    if (s->executed != s->enqueued) return false;
  return true;
}
}  // namespace

void mock_enqueue(hipStream_t st, std::function<void()> f) {
  std::lock_guard<std::mutex> g(g_mu);
  if (!st) {
    if (!g_null) g_null = new_stream();
    st = g_null;
  }
  st->queue.push_back(std::move(f));
  st->enqueued++;
  g_cv.notify_all();
}

// Device time per command, microseconds (0: the device keeps up with the host).
void mock_set_delay(int us) {
  std::lock_guard<std::mutex> g(g_mu);
  g_delay_us = us;
}

// Waits until every stream has executed all its commands (what the
// device-wide synchronizations of HIP wait for).
void mock_drain() {
  std::unique_lock<std::mutex> lk(g_mu);
  g_done_cv.wait(lk, [] { return all_idle(); });
}

static void stop_stream(MockStream *s) {
  {
    std::lock_guard<std::mutex> g(g_mu);
    s->stop = true;
    g_cv.notify_all();
  }
  s->th.join();
  std::lock_guard<std::mutex> g(g_mu);
  g_streams.erase(s);
  delete s;
}

// Stops the stream threads (all commands executed first).
void mock_shutdown() {
  if (g_null) {
    stop_stream(g_null);
    g_null = nullptr;
  }
}

hipError_t hipMalloc(void **p, size_t n) {
  *p = malloc(n ? n : 1);
  return *p ? hipSuccess : hipErrorInvalidValue;
}
// frees synchronize the whole device, as in HIP
hipError_t hipFree(void *p) {
  mock_drain();
  free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned) {
  mock_drain();
  return hipMalloc(p, n);
}
hipError_t hipHostFree(void *p) { return hipFree(p); }
hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind) {
  mock_drain();
  memcpy(dst, src, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind, hipStream_t st) {
  // the copy reads the source when the stream reaches it, as an
  // asynchronous copy from pinned memory does
  mock_enqueue(st, [=] { memcpy(dst, src, n); });
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
hipError_t hipEventRecord(hipEvent_t e, hipStream_t st) {
  {
    std::lock_guard<std::mutex> g(g_mu);
    e->done = false;
  }
  mock_enqueue(st, [e] {
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
  std::lock_guard<std::mutex> g(g_mu);
  *st = new_stream();
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t st) {
  std::unique_lock<std::mutex> lk(g_mu);
  g_done_cv.wait(lk, [st] { return st->executed == st->enqueued; });
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t st) {
  (void)hipStreamSynchronize(st);
  stop_stream(st);
  return hipSuccess;
}
hipError_t hipGetDevice(int *d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "mock error"; }
