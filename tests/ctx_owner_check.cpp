// Host-only check of the context ownership rule of guarded() (objects.h):
// a context serves one thread at a time.  Built by tests/test_ctx_owner.py
// with the host compiler (under ThreadSanitizer when available); it makes no
// HIP call, so it runs without a GPU.
//
//   1. eight threads, each with its own context, make many calls that write
//      the context's error text, kernel-timer totals and counters: every call
//      returns its own status, none is refused;
//   2. two threads on one context: while the first thread's call is running,
//      the second thread's call returns EDX_ERR_BUSY with its text in that
//      thread's own error string, and succeeds once the first call is over;
//   3. a call nested inside another on the same thread is allowed, and the
//      context is free again afterwards.
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <mutex>
#include <thread>
#include <vector>

#include "../realspace_exatdiagonalization_py_amd/csrc/objects.h"

namespace edx {
std::string& tls_error() {
  thread_local std::string s;
  return s;
}
}  // namespace edx

static int failures = 0;
#define CHECK(cond)                                                  \
  do {                                                               \
    if (!(cond)) {                                                   \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
      ++failures;                                                    \
    }                                                                \
  } while (0)

static void own_contexts() {
  constexpr int kThreads = 8, kCalls = 20000;
  std::vector<edx_ctx> ctxs(kThreads);
  std::vector<int> busy(kThreads, 0), bad(kThreads, 0);
  std::vector<std::thread> th;
  for (int t = 0; t < kThreads; ++t)
    th.emplace_back([&, t] {
      edx_ctx* c = &ctxs[t];
      for (int i = 0; i < kCalls; ++i) {
        const int rc = edx::guarded(c, [&] {
          c->timer.ms[i % edx::K_NCLASS] += 1.0;
          c->timer.launches[i % edx::K_NCLASS] += 1;
          c->autotune_group = i & 3;
          if (i % 7 == 0) throw edx::Error{EDX_ERR_INVALID, "call " + std::to_string(i)};
        });
        if (rc == EDX_ERR_BUSY) ++busy[t];
        else if (rc != (i % 7 == 0 ? EDX_ERR_INVALID : EDX_OK)) ++bad[t];
      }
    });
  for (auto& x : th) x.join();
  for (int t = 0; t < kThreads; ++t) {
    CHECK(busy[t] == 0);
    CHECK(bad[t] == 0);
    CHECK(ctxs[t].owner.load() == 0);
    CHECK(ctxs[t].last_error == "call " + std::to_string(kCalls - 1 - (kCalls - 1) % 7));
    int64_t n = 0;
    for (int k = 0; k < edx::K_NCLASS; ++k) n += ctxs[t].timer.launches[k];
    CHECK(n == kCalls);
  }
}

static void shared_context() {
  edx_ctx c;
  std::mutex mu;
  std::condition_variable cv;
  bool inside = false, release = false;
  std::thread a([&] {
    const int rc = edx::guarded(&c, [&] {
      std::unique_lock<std::mutex> lk(mu);
      inside = true;
      cv.notify_all();
      cv.wait(lk, [&] { return release; });
      c.last_error = "first thread";
    });
    CHECK(rc == EDX_OK);
  });
  {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] { return inside; });
  }
  int rc_b = 0;
  std::string text_b;
  std::thread b([&] {
    rc_b = edx::guarded(&c, [&] { c.last_error = "second thread"; });
    text_b = edx::tls_error();
  });
  b.join();
  CHECK(rc_b == EDX_ERR_BUSY);
  CHECK(text_b.find("in use by another thread") != std::string::npos);
  {
    std::lock_guard<std::mutex> lk(mu);
    release = true;
  }
  cv.notify_all();
  a.join();
  CHECK(c.last_error == "first thread");
  CHECK(c.owner.load() == 0);
  std::thread b2([&] { CHECK(edx::guarded(&c, [&] { c.last_error = "second thread"; }) == EDX_OK); });
  b2.join();
  CHECK(c.last_error == "second thread");
}

static void nested_call() {
  edx_ctx c;
  const int rc = edx::guarded(&c, [&] {
    CHECK(edx::guarded(&c, [&] { c.autotune_group = 2; }) == EDX_OK);
    CHECK(c.owner.load() != 0);
  });
  CHECK(rc == EDX_OK);
  CHECK(c.owner.load() == 0 && c.depth == 0);
  CHECK(edx::guarded(nullptr, [] {}) == EDX_OK);
}

int main() {
  own_contexts();
  shared_context();
  nested_call();
  if (failures) {
    std::fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("ctx ownership checks passed\n");
  return 0;
}
