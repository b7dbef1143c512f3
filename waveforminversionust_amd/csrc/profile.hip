// Per-kernel-class HIP-event timing: a pool of event pairs recorded on the
// launching stream around each bracketed launch, resolved and accumulated on
// flush.  Process-wide, for single-threaded callers (the bench and tests).
// Event creation costs host time comparable to a short kernel, so the pool
// is filled when profiling is enabled and events of launches that have
// completed are recycled before new ones are created: the launch loop then
// runs ahead of the device as it does unprofiled.
#include <deque>
#include <mutex>
#include <vector>

#include "profile.h"

namespace wust {

namespace {
struct Pending {
    int cls;
    hipEvent_t a, b;
};
struct Profiler {
    std::mutex mu;
    uint32_t mask = 0;
    std::deque<Pending> pending;
    std::vector<hipEvent_t> pool;
    std::vector<hipEvent_t> open;   // begin events waiting for their end, per class
    double total_ms[PC_COUNT] = {};
    int64_t count[PC_COUNT] = {};
    int64_t units[PC_COUNT] = {};   // work units of the profiled launches (per class; see prof_units)
    int64_t announced[PC_COUNT] = {};  // units announced for a class's next launch
    int64_t calls[PC_COUNT] = {};    // launches of a profiled class seen (for sampling)
    uint32_t stride = 1;             // time one launch in `stride` of each profiled class
    Profiler() : open(PC_COUNT, nullptr) {}
    // resolve one pending pair (blocking when `wait`); false if it has not completed
    bool resolve_front(bool wait) {
        Pending &q = pending.front();
        if (wait) {
            WUST_HIP_CHECK(hipEventSynchronize(q.b));
        } else if (hipEventQuery(q.b) != hipSuccess) {
            return false;
        }
        float ms = 0;
        WUST_HIP_CHECK(hipEventElapsedTime(&ms, q.a, q.b));
        total_ms[q.cls] += ms;
        count[q.cls] += 1;
        pool.push_back(q.a);
        pool.push_back(q.b);
        pending.pop_front();
        return true;
    }
    void grow(int n) {
        for (int i = 0; i < n; ++i) {
            hipEvent_t e;
            WUST_HIP_CHECK(hipEventCreate(&e));
            pool.push_back(e);
        }
    }
    hipEvent_t get() {
        if (pool.empty()) {
            // recycle the events of launches that have completed (oldest first)
            for (int i = 0; i < 256 && !pending.empty(); ++i)
                if (!resolve_front(false)) break;
        }
        if (pool.empty()) grow(64);
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
};
Profiler &P() {
    static Profiler p;
    return p;
}
}  // namespace

bool prof_on(int cls) { return (P().mask >> cls) & 1u; }

void prof_record(int cls, bool begin, hipStream_t s) {
    Profiler &p = P();
    std::lock_guard<std::mutex> lk(p.mu);
    hipEvent_t e = p.get();
    WUST_HIP_CHECK(hipEventRecord(e, s));
    if (begin) {
        p.open[cls] = e;
    } else if (p.open[cls]) {
        p.pending.push_back({cls, p.open[cls], e});
        p.open[cls] = nullptr;
    } else {
        p.pool.push_back(e);
    }
}

hipEvent_t prof_event() {
    Profiler &p = P();
    std::lock_guard<std::mutex> lk(p.mu);
    return p.get();
}

void prof_push(int cls, hipEvent_t a, hipEvent_t b) {
    Profiler &p = P();
    std::lock_guard<std::mutex> lk(p.mu);
    p.pending.push_back({cls, a, b});
}

void prof_units(int cls, int64_t n) {
    if (!prof_on(cls)) return;
    Profiler &p = P();
    std::lock_guard<std::mutex> lk(p.mu);
    p.announced[cls] = n;
}

bool prof_take(int cls) {
    if (!prof_on(cls)) return false;
    Profiler &p = P();
    std::lock_guard<std::mutex> lk(p.mu);
    const bool take = (p.calls[cls]++ % p.stride) == 0;
    if (take) p.units[cls] += p.announced[cls];
    p.announced[cls] = 0;
    return take;
}

void prof_flush() {
    Profiler &p = P();
    std::lock_guard<std::mutex> lk(p.mu);
    while (!p.pending.empty()) p.resolve_front(true);
}

}  // namespace wust

extern "C" {

int wust_profile_enable(uint32_t class_mask) {
    wust::Profiler &p = wust::P();
    try {
        std::lock_guard<std::mutex> lk(p.mu);
        if (class_mask && p.pool.size() < 4096) p.grow(4096 - (int)p.pool.size());
    } catch (const wust::Error &e) {
        return e.status;
    }
    p.mask = class_mask;
    return WUST_OK;
}

int wust_profile_read(int cls, int64_t *count, double *total_ms) {
    if (cls < 0 || cls >= wust::PC_COUNT || !count || !total_ms) return WUST_INVALID_INPUT;
    try {
        wust::prof_flush();
    } catch (const wust::Error &e) {
        return e.status;
    }
    *count = wust::P().count[cls];
    *total_ms = wust::P().total_ms[cls];
    return WUST_OK;
}

int wust_profile_stride(uint32_t stride) {
    if (stride < 1) return WUST_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(wust::P().mu);
    wust::P().stride = stride;
    for (int c = 0; c < wust::PC_COUNT; ++c) wust::P().calls[c] = 0;
    return WUST_OK;
}

int wust_profile_units(int cls, int64_t *units) {
    if (cls < 0 || cls >= wust::PC_COUNT || !units) return WUST_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(wust::P().mu);
    *units = wust::P().units[cls];
    return WUST_OK;
}

int wust_profile_reset(void) {
    try {
        wust::prof_flush();
    } catch (const wust::Error &e) {
        return e.status;
    }
    for (int c = 0; c < wust::PC_COUNT; ++c) {
        wust::P().count[c] = 0;
        wust::P().total_ms[c] = 0;
        wust::P().units[c] = 0;
        wust::P().announced[c] = 0;
        wust::P().calls[c] = 0;
    }
    return WUST_OK;
}

}  // extern "C"
