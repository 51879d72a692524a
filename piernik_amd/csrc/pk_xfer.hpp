// pk_xfer.hpp -- host side of the transfers between the caller's (pageable)
// arrays and the device: a ring of pinned chunks, filled or drained by a few
// worker threads while the DMA engine moves the previous chunk, so the host
// copy and the PCIe transfer overlap instead of running one after the other
// inside a single pageable hipMemcpy.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace pk_xfer {

// memcpy split over the caller and a fixed set of worker threads.
class CopyPool {
public:
    ~CopyPool() { stop(); }

    void start(int nthreads)
    {
        if (!workers_.empty() || nthreads <= 0) return;
        quit_ = false;
        for (int i = 0; i < nthreads; i++) workers_.emplace_back([this]() { loop(); });
    }

    void stop()
    {
        {
            std::lock_guard<std::mutex> g(mu_);
            quit_ = true;
        }
        cv_.notify_all();
        for (auto &t : workers_) t.join();
        workers_.clear();
    }

    // dst[0..n) = src[0..n); returns when every part is copied
    void copy(void *dst, const void *src, size_t n)
    {
        const size_t kMinPart = 1u << 20;
        const int parts = (int)std::min<size_t>(workers_.size() + 1, std::max<size_t>(1, n / kMinPart));
        if (parts <= 1) {
            std::memcpy(dst, src, n);
            return;
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            dst_ = static_cast<char *>(dst);
            src_ = static_cast<const char *>(src);
            n_ = n;
            parts_ = parts;
            next_ = 0;
            done_ = 0;
            gen_++;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> l(mu_);
        done_cv_.wait(l, [this]() { return done_ == parts_; });
    }

private:
    void loop()
    {
        unsigned seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> l(mu_);
                cv_.wait(l, [&]() { return quit_ || gen_ != seen; });
                if (quit_) return;
                seen = gen_;
            }
            work();
        }
    }

    // take parts until none is left
    void work()
    {
        for (;;) {
            int i;
            char *d;
            const char *s;
            size_t n;
            int parts;
            {
                std::lock_guard<std::mutex> g(mu_);
                if (next_ >= parts_) return;
                i = next_++;
                d = dst_;
                s = src_;
                n = n_;
                parts = parts_;
            }
            const size_t per = (n + parts - 1) / parts;
            const size_t b = std::min(n, (size_t)i * per), e = std::min(n, b + per);
            if (e > b) std::memcpy(d + b, s + b, e - b);
            {
                std::lock_guard<std::mutex> g(mu_);
                done_++;
                if (done_ == parts_) done_cv_.notify_all();
            }
        }
    }

    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t n_ = 0;
    int parts_ = 0, next_ = 0, done_ = 0;
    unsigned gen_ = 0;
    bool quit_ = false;
};

// Pinned chunks and one event per chunk (the last DMA that used it).  Every
// transfer of one Ring goes on the same stream, so a DMA into a slot is
// ordered after the earlier DMA out of it; the host waits on the slot's
// event before it writes or reads the slot itself.
struct Ring {
    static constexpr int kSlots = 4;
    static constexpr size_t kChunk = 32u << 20;
    char *slot[kSlots] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t ev[kSlots] = {nullptr, nullptr, nullptr, nullptr};
    CopyPool pool;

    hipError_t ensure()
    {
        if (slot[0]) return hipSuccess;
        for (int i = 0; i < kSlots; i++) {
            hipError_t e = hipHostMalloc((void **)&slot[i], kChunk, hipHostMallocDefault);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
            if (e != hipSuccess) {
                release();
                return e;
            }
        }
        const unsigned hc = std::thread::hardware_concurrency();
        pool.start((int)std::min(7u, hc > 1 ? hc - 1 : 0u));
        return hipSuccess;
    }

    void release()
    {
        pool.stop();
        for (int i = 0; i < kSlots; i++) {
            if (ev[i]) (void)hipEventDestroy(ev[i]);
            if (slot[i]) (void)hipHostFree(slot[i]);
            ev[i] = nullptr;
            slot[i] = nullptr;
        }
    }

    // device dst[0..n) = host src[0..n) on stream st; returns once every
    // chunk is queued (the DMA may still run: order later work on st)
    hipError_t h2d(void *dst, const void *src, size_t n, hipStream_t st)
    {
        hipError_t e = ensure();
        if (e != hipSuccess) return e;
        int k = 0;
        for (size_t off = 0; off < n; off += kChunk, k++) {
            const size_t len = std::min(kChunk, n - off);
            const int s = k % kSlots;
            // the last DMA out of this slot (of this call or an earlier one) is done
            if ((e = hipEventSynchronize(ev[s])) != hipSuccess) return e;
            pool.copy(slot[s], static_cast<const char *>(src) + off, len);
            if ((e = hipMemcpyAsync(static_cast<char *>(dst) + off, slot[s], len, hipMemcpyHostToDevice, st)) !=
                hipSuccess)
                return e;
            if ((e = hipEventRecord(ev[s], st)) != hipSuccess) return e;
        }
        return hipSuccess;
    }

    // host dst[0..n) = device src[0..n) after the work queued on st; returns
    // when the host copy is complete
    hipError_t d2h(void *dst, const void *src, size_t n, hipStream_t st)
    {
        hipError_t e = ensure();
        if (e != hipSuccess) return e;
        const size_t nch = (n + kChunk - 1) / kChunk;
        auto issue = [&](size_t k) -> hipError_t {
            const size_t off = k * kChunk, len = std::min(kChunk, n - off);
            const int s = (int)(k % kSlots);
            hipError_t r = hipMemcpyAsync(slot[s], static_cast<const char *>(src) + off, len,
                                          hipMemcpyDeviceToHost, st);
            if (r == hipSuccess) r = hipEventRecord(ev[s], st);
            return r;
        };
        for (size_t k = 0; k < std::min<size_t>(nch, kSlots); k++)
            if ((e = issue(k)) != hipSuccess) return e;
        for (size_t k = 0; k < nch; k++) {
            const int s = (int)(k % kSlots);
            if ((e = hipEventSynchronize(ev[s])) != hipSuccess) return e;
            const size_t off = k * kChunk, len = std::min(kChunk, n - off);
            pool.copy(static_cast<char *>(dst) + off, slot[s], len);
            if (k + kSlots < nch && (e = issue(k + kSlots)) != hipSuccess) return e;
        }
        return hipSuccess;
    }
};

}  // namespace pk_xfer
