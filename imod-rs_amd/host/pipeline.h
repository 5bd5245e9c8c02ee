// Two-buffer host pipeline shared by the drop-in front ends (`tilt`,
// `ctfphaseflip`): file reads overlap device uploads, and device read-backs
// overlap file writes.
#pragma once

#include <condition_variable>
#include <functional>
#include <mutex>
#include <string>
#include <thread>

namespace tiltgpu_host {

// Two-buffer pipeline: produce(k, buf) runs on a worker thread, consume(k, buf)
// on the caller's, chunk k in buffer k % 2.  Stops at the first failure (its
// message in *err) and joins the worker before returning.
inline bool pipeline2(int chunks, float* const bufs[2],
                      const std::function<bool(int, float*, std::string*)>& produce,
                      const std::function<bool(int, float*, std::string*)>& consume,
                      std::string* err) {
    std::mutex m;
    std::condition_variable cv;
    int full[2] = {0, 0};
    bool failed = false;
    std::string werr;
    std::thread worker([&] {
        for (int k = 0; k < chunks; ++k) {
            const int b = k & 1;
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return full[b] == 0 || failed; });
                if (failed) return;
            }
            std::string e;
            const bool ok = produce(k, bufs[b], &e);
            std::lock_guard<std::mutex> lk(m);
            if (!ok) {
                failed = true;
                werr = e;
            } else {
                full[b] = 1;
            }
            cv.notify_all();
            if (!ok) return;
        }
    });
    bool ok = true;
    for (int k = 0; k < chunks && ok; ++k) {
        const int b = k & 1;
        {
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [&] { return full[b] == 1 || failed; });
            if (failed) {
                ok = false;
                break;
            }
        }
        std::string e;
        ok = consume(k, bufs[b], &e);
        std::lock_guard<std::mutex> lk(m);
        if (!ok) {
            failed = true;
            werr = e;
        }
        full[b] = 0;
        cv.notify_all();
    }
    worker.join();
    if (failed || !ok) {
        *err = werr;
        return false;
    }
    return true;
}

}  // namespace tiltgpu_host
