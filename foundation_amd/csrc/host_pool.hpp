// host_pool.hpp — worker pools of the host-side builders, and the guard of
// the host-only entry points.  No exception leaves a worker thread (that
// would terminate the process) and none crosses the C ABI.
#pragma once
#include <atomic>
#include <cstdio>
#include <new>
#include <stdexcept>
#include <thread>
#include <vector>

#include "../../include/fnd_pt.h"

namespace fnd {

// Runs worker() on the calling thread and on up to n - 1 more threads, and
// waits for all of them.  Workers share their work through a counter of their
// own, so a thread that cannot be started only leaves its share to the
// others.  An exception inside a worker ends that worker; after the join the
// first one is raised again on the calling thread as std::bad_alloc (host
// memory ran out) or std::runtime_error (anything else).
template <class W>
void run_pool(unsigned n, W&& worker)
{
    std::atomic<int> failure{0};  // 0 none, 1 out of memory, 2 other
    auto body = [&]() noexcept {
        try {
            worker();
        } catch (const std::bad_alloc&) {
            int none = 0;
            failure.compare_exchange_strong(none, 1);
        } catch (...) {
            int none = 0;
            failure.compare_exchange_strong(none, 2);
        }
    };
    std::vector<std::thread> pool;
    try {
        pool.reserve(n > 1 ? n - 1 : 0);
        for (unsigned i = 1; i < n; ++i) pool.emplace_back(body);
    } catch (...) {
        // fewer threads than asked: the ones started and this one share the work
    }
    body();
    for (auto& th : pool) th.join();
    if (failure.load() == 1) throw std::bad_alloc();
    if (failure.load() == 2) throw std::runtime_error("a builder worker failed");
}

// The body of a host-only entry point that reports failures as text in the
// caller's err buffer (may be null): a host allocation failure becomes
// FND_E_OOM_HOST, anything else FND_E_INVALID_ARG with the exception's text.
template <class F>
fnd_status guard_host(char* err, size_t err_len, F&& body) noexcept
{
    auto say = [&](const char* msg) {
        if (err && err_len) std::snprintf(err, err_len, "%s", msg);
    };
    try {
        return body();
    } catch (const std::bad_alloc&) {
        say("out of host memory");
        return FND_E_OOM_HOST;
    } catch (const std::length_error&) {
        say("out of host memory (container size limit)");
        return FND_E_OOM_HOST;
    } catch (const std::exception& e) {
        say(e.what());
        return FND_E_INVALID_ARG;
    } catch (...) {
        say("internal error");
        return FND_E_INVALID_ARG;
    }
}

}  // namespace fnd
