// Per-kernel-class HIP-event timing: a pool of event pairs recorded on the
// launching stream around each bracketed launch, resolved and accumulated on
// flush.  Process-wide, for single-threaded callers (the bench and tests).
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
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    std::vector<hipEvent_t> open;   // begin events waiting for their end, per class
    double total_ms[PC_COUNT] = {};
    int64_t count[PC_COUNT] = {};
    Profiler() : open(PC_COUNT, nullptr) {}
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        WUST_HIP_CHECK(hipEventCreate(&e));
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

void prof_flush() {
    Profiler &p = P();
    std::lock_guard<std::mutex> lk(p.mu);
    for (auto &q : p.pending) {
        WUST_HIP_CHECK(hipEventSynchronize(q.b));
        float ms = 0;
        WUST_HIP_CHECK(hipEventElapsedTime(&ms, q.a, q.b));
        p.total_ms[q.cls] += ms;
        p.count[q.cls] += 1;
        p.pool.push_back(q.a);
        p.pool.push_back(q.b);
    }
    p.pending.clear();
}

}  // namespace wust

extern "C" {

int wust_profile_enable(uint32_t class_mask) {
    wust::P().mask = class_mask;
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

int wust_profile_reset(void) {
    try {
        wust::prof_flush();
    } catch (const wust::Error &e) {
        return e.status;
    }
    for (int c = 0; c < wust::PC_COUNT; ++c) {
        wust::P().count[c] = 0;
        wust::P().total_ms[c] = 0;
    }
    return WUST_OK;
}

}  // extern "C"
