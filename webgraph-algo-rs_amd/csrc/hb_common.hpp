// Shared host-side helpers of the HyperBall library: error plumbing across
// the C ABI and a small thread pool for host preprocessing.
#pragma once

#include <algorithm>
#include <cstdint>
#include <exception>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "hyperball.h"

namespace hb {

struct Error : std::runtime_error {
  hb_status code;
  Error(hb_status c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

[[noreturn]] inline void fail(hb_status code, const std::string &msg) { throw Error(code, msg); }

void set_last_error(const std::string &msg);

// Runs f, mapping exceptions to status codes and recording the message.
template <class F>
hb_status guard(F &&f) {
  try {
    f();
    return HB_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return HB_E_OOM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return HB_E_INVALID_ARG;
  } catch (...) {
    set_last_error("unknown error");
    return HB_E_INVALID_ARG;
  }
}

inline unsigned host_threads() {
  unsigned t = std::thread::hardware_concurrency();
  return std::max(1u, std::min(t ? t : 1u, 64u));
}

// Static block partition of [0, n) over the host threads.
template <class F>
void parallel_for(uint64_t n, F &&f) {
  unsigned t = host_threads();
  if (n < 4096 || t == 1) {
    f(uint64_t(0), n, 0u);
    return;
  }
  std::vector<std::thread> pool;
  std::mutex err_mu;
  std::exception_ptr err;  // the first failure, rethrown after the join
  uint64_t chunk = (n + t - 1) / t;
  for (unsigned i = 0; i < t; i++) {
    uint64_t b = std::min(n, i * chunk), e = std::min(n, b + chunk);
    pool.emplace_back([&, b, e, i] {
      try {
        f(b, e, i);
      } catch (...) {
        std::lock_guard<std::mutex> l(err_mu);
        if (!err) err = std::current_exception();
      }
    });
  }
  for (auto &th : pool) th.join();
  if (err) std::rethrow_exception(err);
}

}  // namespace hb
