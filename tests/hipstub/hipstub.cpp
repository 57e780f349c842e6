// Test-only HIP runtime stand-in: streams are host worker threads, device memory is host memory, graphs are
// recorded operation lists. It follows the HIP runtime API documentation where the library relies on it and
// records every use that the documentation (or the stream-capture rules of the runtime) does not allow:
//   - hipFree / hipHostFree perform an implicit hipDeviceSynchronize (hip_runtime_api.h): a device-wide
//     synchronisation while a stream of another thread is being captured synchronises a stream whose
//     capture is in progress, and is recorded;
//   - in the thread that captures, calls that synchronise or allocate (hipMalloc, hipFree, hipHostMalloc,
//     hipHostFree, stream/event/device synchronisation) are recorded;
//   - a capture whose forked streams were not joined back before hipStreamEndCapture is recorded;
//   - copies whose device side is not (or no longer) inside a live hipMalloc allocation are recorded, at
//     the call and again when the copy executes.
// Pageable host memory follows the documented semantics of hipMemcpyAsync: host-to-device copies read the
// source before returning (staging), device-to-host copies complete before returning. Allocations are
// filled with NaN bytes so that reads of device memory that nothing initialised show up in the results.
#include "hipstub.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <shared_mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

struct Capture;

}  // namespace

struct ihipStream_t {
  std::mutex m;
  std::condition_variable cv;
  std::deque<std::function<void()>> q;
  uint64_t enq = 0, done = 0;
  bool stop = false;
  std::thread th;
  Capture* cap = nullptr;  // guarded by g_mu
  uint64_t cap_ops = 0;    // operations captured from this stream in the current capture (g_mu)
};

struct ihipEvent_t {
  std::mutex m;
  std::condition_variable cv;
  uint64_t gen_rec = 0, gen_done = 0;
  double t = 0;
  // recorded during a capture (g_mu)
  Capture* cap = nullptr;
  ihipStream_t* cap_stream = nullptr;
  uint64_t cap_seq = 0;
};

struct ihipGraph {
  std::shared_ptr<std::vector<std::function<void()>>> ops;
};
struct hipGraphExec {
  std::shared_ptr<std::vector<std::function<void()>>> ops;
};

namespace {

struct Capture {
  std::shared_ptr<std::vector<std::function<void()>>> ops = std::make_shared<std::vector<std::function<void()>>>();
  std::thread::id owner;
  ihipStream_t* origin = nullptr;
  std::set<ihipStream_t*> members, unjoined;
  std::set<ihipEvent_t*> events;
};

struct Alloc {
  size_t size;
  bool pinned;
};

std::mutex g_mu;
std::shared_mutex g_life;  // device_sync (shared) keeps the streams it drains alive against hipStreamDestroy
std::set<ihipStream_t*> g_streams;
std::set<Capture*> g_captures;
std::map<uintptr_t, Alloc> g_allocs;
std::vector<std::string> g_viol;
int g_max_caps = 0;

void violation_locked(const std::string& s) {
  if (g_viol.size() < 1000) g_viol.push_back(s);
  std::fprintf(stderr, "[hipstub] violation: %s\n", s.c_str());
}
void violation(const std::string& s) {
  std::lock_guard<std::mutex> lk(g_mu);
  violation_locked(s);
}

// live allocation holding [p, p + n): 1 = device, 2 = pinned host, 0 = none (g_mu held)
int find_locked(const void* p, size_t n) {
  const uintptr_t a = (uintptr_t)p;
  auto it = g_allocs.upper_bound(a);
  if (it == g_allocs.begin()) return 0;
  --it;
  if (a >= it->first && a + n <= it->first + it->second.size) return it->second.pinned ? 2 : 1;
  return 0;
}

// capture owned by the calling thread, if any (g_mu held)
Capture* own_capture_locked() {
  for (Capture* c : g_captures)
    if (c->owner == std::this_thread::get_id()) return c;
  return nullptr;
}

// an allocating or synchronising call: not allowed in a thread that is capturing
void unsafe_call(const char* what) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (own_capture_locked()) violation_locked(std::string(what) + " called by the capturing thread during a stream capture");
}

void worker(ihipStream_t* s) {
  std::unique_lock<std::mutex> lk(s->m);
  while (true) {
    s->cv.wait(lk, [&] { return s->stop || !s->q.empty(); });
    if (s->q.empty()) return;
    auto fn = std::move(s->q.front());
    s->q.pop_front();
    lk.unlock();
    fn();
    lk.lock();
    s->done++;
    s->cv.notify_all();
  }
}

void push(ihipStream_t* s, std::function<void()> fn) {
  std::lock_guard<std::mutex> lk(s->m);
  s->q.push_back(std::move(fn));
  s->enq++;
  s->cv.notify_all();
}

void drain(ihipStream_t* s) {
  std::unique_lock<std::mutex> lk(s->m);
  const uint64_t target = s->enq;
  s->cv.wait(lk, [&] { return s->done >= target; });
}

// device-wide synchronisation (hipDeviceSynchronize, and implicitly hipFree / hipHostFree)
void device_sync(const char* what) {
  std::shared_lock<std::shared_mutex> life(g_life);
  std::vector<ihipStream_t*> ss;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (own_capture_locked()) violation_locked(std::string(what) + " called by the capturing thread during a stream capture");
    for (Capture* c : g_captures)
      if (c->owner != std::this_thread::get_id())
        violation_locked(std::string(what) +
                         " synchronises the device (implicit hipDeviceSynchronize) while another thread captures a stream");
    for (ihipStream_t* s : g_streams)
      if (!s->cap) ss.push_back(s);
  }
  for (ihipStream_t* s : ss) drain(s);
}

thread_local hipError_t t_last = hipSuccess;

}  // namespace

namespace hipstub {
void enqueue(hipStream_t s, std::function<void()> fn) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_streams.count(s)) {
      violation_locked("launch on a stream that does not exist");
      return;
    }
    if (s->cap) {
      s->cap->ops->push_back(std::move(fn));
      s->cap_ops++;
      return;
    }
  }
  push(s, std::move(fn));
}
}  // namespace hipstub

extern "C" {

int hipstub_violation_count(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  return (int)g_viol.size();
}
const char* hipstub_violation(int i) {
  std::lock_guard<std::mutex> lk(g_mu);
  return (i >= 0 && i < (int)g_viol.size()) ? g_viol[i].c_str() : "";
}
long long hipstub_live_device_bytes(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  long long n = 0;
  for (auto& kv : g_allocs)
    if (!kv.second.pinned) n += (long long)kv.second.size;
  return n;
}
int hipstub_max_concurrent_captures(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_max_caps;
}

const char* hipGetErrorString(hipError_t e) {
  switch (e) {
  case hipSuccess: return "hipSuccess";
  case hipErrorInvalidValue: return "hipErrorInvalidValue";
  case hipErrorOutOfMemory: return "hipErrorOutOfMemory";
  case hipErrorInvalidDevice: return "hipErrorInvalidDevice";
  case hipErrorInvalidHandle: return "hipErrorInvalidHandle";
  case hipErrorStreamCaptureUnsupported: return "hipErrorStreamCaptureUnsupported";
  case hipErrorStreamCaptureUnjoined: return "hipErrorStreamCaptureUnjoined";
  default: return "hipError (stub)";
  }
}
hipError_t hipGetLastError(void) {
  const hipError_t e = t_last;
  t_last = hipSuccess;
  return e;
}
hipError_t hipGetDeviceCount(int* n) {
  *n = 1;
  return hipSuccess;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidDevice; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int d) {
  if (d != 0) return hipErrorInvalidDevice;
  std::memset(p, 0, sizeof(*p));
  std::snprintf(p->name, sizeof(p->name), "host stub");
  std::snprintf(p->gcnArchName, sizeof(p->gcnArchName), "cpu");
  return hipSuccess;
}
hipError_t hipDeviceSynchronize(void) {
  device_sync("hipDeviceSynchronize");
  return hipSuccess;
}

hipError_t hipMalloc(void** ptr, size_t size) {
  unsafe_call("hipMalloc");
  if (size == 0) {
    *ptr = nullptr;
    return hipSuccess;
  }
  void* p = std::malloc(size);
  if (!p) return hipErrorOutOfMemory;
  std::memset(p, 0xff, size);  // NaN in every float and double
  std::lock_guard<std::mutex> lk(g_mu);
  g_allocs[(uintptr_t)p] = Alloc{size, false};
  *ptr = p;
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  device_sync("hipFree");
  if (!p) return hipSuccess;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_allocs.find((uintptr_t)p);
    if (it == g_allocs.end() || it->second.pinned) {
      violation_locked("hipFree of a pointer that hipMalloc did not return");
      return hipErrorInvalidValue;
    }
    g_allocs.erase(it);
  }
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** ptr, size_t size, unsigned int) {
  unsafe_call("hipHostMalloc");
  void* p = std::malloc(size ? size : 1);
  if (!p) return hipErrorOutOfMemory;
  std::lock_guard<std::mutex> lk(g_mu);
  g_allocs[(uintptr_t)p] = Alloc{size ? size : 1, true};
  *ptr = p;
  return hipSuccess;
}
hipError_t hipHostFree(void* p) {
  device_sync("hipHostFree");
  if (!p) return hipSuccess;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_allocs.find((uintptr_t)p);
    if (it == g_allocs.end() || !it->second.pinned) {
      violation_locked("hipHostFree of a pointer that hipHostMalloc did not return");
      return hipErrorInvalidValue;
    }
    g_allocs.erase(it);
  }
  std::free(p);
  return hipSuccess;
}

hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  const bool d_dev = kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToDevice;
  const bool s_dev = kind == hipMemcpyDeviceToHost || kind == hipMemcpyDeviceToDevice;
  bool pageable_src = false, pageable_dst = false, capturing = false;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_streams.count(s)) {
      violation_locked("hipMemcpyAsync on a stream that does not exist");
      return hipErrorInvalidHandle;
    }
    capturing = s->cap != nullptr;
    if (d_dev && find_locked(dst, bytes) != 1) violation_locked("hipMemcpyAsync: destination outside every device allocation");
    if (s_dev && find_locked(src, bytes) != 1) violation_locked("hipMemcpyAsync: source outside every device allocation");
    if (!s_dev && find_locked(src, bytes) != 2) pageable_src = true;
    if (!d_dev && find_locked(dst, bytes) != 2) pageable_dst = true;
    if (capturing && (pageable_src || pageable_dst))
      violation_locked("hipMemcpyAsync with pageable host memory during a stream capture");
  }
  auto check = [=](const char* when) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (d_dev && find_locked(dst, bytes) != 1)
      violation_locked(std::string("copy ") + when + ": destination no longer inside a device allocation");
    if (s_dev && find_locked(src, bytes) != 1)
      violation_locked(std::string("copy ") + when + ": source no longer inside a device allocation");
    if (!s_dev && !pageable_src && find_locked(src, bytes) != 2)
      violation_locked(std::string("copy ") + when + ": pinned source freed");
    if (!d_dev && !pageable_dst && find_locked(dst, bytes) != 2)
      violation_locked(std::string("copy ") + when + ": pinned destination freed");
  };
  if (pageable_src && !capturing) {  // staged: the caller may reuse its buffer on return
    auto staged = std::make_shared<std::vector<char>>((const char*)src, (const char*)src + bytes);
    hipstub::enqueue(s, [=]() {
      check("at execution");
      std::memcpy(dst, staged->data(), bytes);
    });
    return hipSuccess;
  }
  hipstub::enqueue(s, [=]() {
    check("at execution");
    std::memcpy(dst, src, bytes);
  });
  if (pageable_dst && !capturing) drain(s);  // returns when the copy has completed
  return hipSuccess;
}

hipError_t hipStreamCreateWithFlags(hipStream_t* out, unsigned int) {
  auto* s = new ihipStream_t();
  s->th = std::thread(worker, s);
  std::lock_guard<std::mutex> lk(g_mu);
  g_streams.insert(s);
  *out = s;
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_streams.count(s)) return hipErrorInvalidHandle;
    if (s->cap) violation_locked("hipStreamDestroy of a stream that is being captured");
  }
  drain(s);
  {
    std::unique_lock<std::shared_mutex> life(g_life);
    std::lock_guard<std::mutex> lk(g_mu);
    g_streams.erase(s);
  }
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->stop = true;
    s->cv.notify_all();
  }
  s->th.join();
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  unsafe_call("hipStreamSynchronize");
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_streams.count(s)) return hipErrorInvalidHandle;
    if (s->cap) {
      violation_locked("hipStreamSynchronize of a stream that is being captured");
      return hipErrorStreamCaptureUnsupported;
    }
  }
  drain(s);
  return hipSuccess;
}

hipError_t hipEventCreate(hipEvent_t* e) {
  *e = new ihipEvent_t();
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned int) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  if (!e) return hipErrorInvalidHandle;
  {
    std::unique_lock<std::mutex> lk(e->m);  // pending records complete first
    e->cv.wait(lk, [&] { return e->gen_done >= e->gen_rec; });
  }
  {
    std::lock_guard<std::mutex> lk(g_mu);
    for (Capture* c : g_captures) c->events.erase(e);
  }
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  uint64_t gen = 0;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_streams.count(s)) return hipErrorInvalidHandle;
    if (s->cap) {
      e->cap = s->cap;
      e->cap_stream = s;
      e->cap_seq = s->cap_ops;
      s->cap->events.insert(e);
      return hipSuccess;
    }
    e->cap = nullptr;
    std::lock_guard<std::mutex> lk2(e->m);
    gen = ++e->gen_rec;
  }
  push(s, [e, gen]() {
    std::lock_guard<std::mutex> lk(e->m);
    e->t = now_s();
    e->gen_done = std::max(e->gen_done, gen);
    e->cv.notify_all();
  });
  return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned int) {
  uint64_t gen = 0;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_streams.count(s)) return hipErrorInvalidHandle;
    if (e->cap) {
      Capture* c = e->cap;
      if (!s->cap) {  // fork: s joins the capture
        s->cap = c;
        s->cap_ops = 0;
        c->members.insert(s);
        c->unjoined.insert(s);
      } else if (s->cap != c) {
        violation_locked("hipStreamWaitEvent joins two different captures");
        return hipErrorStreamCaptureUnsupported;
      } else if (s == c->origin && e->cap_stream != s && e->cap_stream->cap_ops == e->cap_seq) {
        c->unjoined.erase(e->cap_stream);  // join: the origin waits for the forked stream's last operation
      }
      return hipSuccess;
    }
    if (s->cap) {
      violation_locked("hipStreamWaitEvent in a capture on an event recorded outside it");
      return hipErrorStreamCaptureUnsupported;
    }
    std::lock_guard<std::mutex> lk2(e->m);
    gen = e->gen_rec;
  }
  if (gen == 0) return hipSuccess;
  push(s, [e, gen]() {
    std::unique_lock<std::mutex> lk(e->m);
    e->cv.wait(lk, [&] { return e->gen_done >= gen; });
  });
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
  unsafe_call("hipEventSynchronize");
  std::unique_lock<std::mutex> lk(e->m);
  const uint64_t gen = e->gen_rec;
  e->cv.wait(lk, [&] { return e->gen_done >= gen; });
  return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  std::lock_guard<std::mutex> la(a->m);
  if (a != b) {
    std::lock_guard<std::mutex> lb(b->m);
    if (a->gen_done < a->gen_rec || b->gen_done < b->gen_rec || a->gen_rec == 0 || b->gen_rec == 0) return hipErrorNotReady;
    *ms = (float)((b->t - a->t) * 1e3);
    return hipSuccess;
  }
  *ms = 0.f;
  return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_streams.count(s)) return hipErrorInvalidHandle;
  if (s->cap) return hipErrorStreamCaptureUnsupported;
  auto* c = new Capture();
  c->owner = std::this_thread::get_id();
  c->origin = s;
  c->members.insert(s);
  s->cap = c;
  s->cap_ops = 0;
  g_captures.insert(c);
  g_max_caps = std::max(g_max_caps, (int)g_captures.size());
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  std::lock_guard<std::mutex> lk(g_mu);
  Capture* c = s->cap;
  if (!c || c->origin != s) {
    violation_locked("hipStreamEndCapture on a stream that did not begin the capture");
    return hipErrorStreamCaptureUnsupported;
  }
  hipError_t rc = hipSuccess;
  if (!c->unjoined.empty()) {
    violation_locked("hipStreamEndCapture with forked streams not joined back");
    rc = hipErrorStreamCaptureUnjoined;
  }
  for (ihipStream_t* m : c->members) {
    m->cap = nullptr;
    m->cap_ops = 0;
  }
  for (ihipEvent_t* e : c->events) e->cap = nullptr;
  g_captures.erase(c);
  if (rc == hipSuccess) {
    auto* gr = new ihipGraph();
    gr->ops = c->ops;
    *g = gr;
  } else {
    *g = nullptr;
  }
  delete c;
  return rc;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* exec, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  if (!g) return hipErrorInvalidValue;
  auto* x = new hipGraphExec();
  x->ops = std::make_shared<std::vector<std::function<void()>>>(*g->ops);
  *exec = x;
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t x) {
  delete x;  // pending launches hold their own reference to the operation list
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s) {
  if (!x) return hipErrorInvalidValue;
  auto ops = x->ops;
  hipstub::enqueue(s, [ops]() {
    for (auto& f : *ops) f();
  });
  return hipSuccess;
}

}  // extern "C"
