// Per-kernel-class HIP-event timing (tooling; off unless enabled).
#pragma once
#include <hip/hip_ext.h>

#include "wust_common.h"

namespace wust {

enum ProfClass {
    PC_ASSEMBLY = 0,
    PC_SCHUR,
    PC_GJ_PANEL,
    PC_SWEEP_OPERAND,
    PC_SWEEP_GEMM,
    PC_RESIDUAL,
    PC_SOURCE_FIT,
    PC_GRADIENT,
    PC_PLAIN_FACTOR_STEP,
    PC_PLAIN_SOLVE_STEP,
    PC_NS_GEMM,
    PC_COUNT
};

bool prof_on(int cls);
void prof_record(int cls, bool begin, hipStream_t s);
void prof_flush();
hipEvent_t prof_event();
void prof_push(int cls, hipEvent_t a, hipEvent_t b);
// announce the work units (the class's own unit, e.g. chain steps of a sweep launch) of the
// class's next launch; they are counted if that launch is timed
void prof_units(int cls, int64_t n);
// whether this launch of cls is timed: the class is profiled and the launch is the first of each
// `stride` (wust_profile_stride) since the last reset
bool prof_take(int cls);

// Launch a kernel; when its class is profiled (and this launch is sampled), the
// dispatch packet itself timestamps start and end (hipExtLaunchKernelGGL
// events), so the timing adds no packets to the stream.  The completion signal
// of such a launch leaves the device idle for several microseconds before the
// next one, which is why a timed run samples one launch in many.
template <typename K, typename... Args>
void launch_k(int cls, K kernel, dim3 grid, dim3 block, uint32_t shmem, hipStream_t s, Args... args) {
    if (prof_take(cls)) {
        hipEvent_t a = prof_event(), b = prof_event();
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, a, b, 0, args...);
        prof_push(cls, a, b);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
    }
    WUST_LAUNCH_CHECK();
}

// RAII bracket around one launch (or a fixed group of launches) of a class.
struct ProfScope {
    int cls;
    hipStream_t s;
    bool on;
    ProfScope(int c, hipStream_t st) : cls(c), s(st), on(prof_on(c)) {
        if (on) prof_record(cls, true, s);
    }
    ~ProfScope() {
        if (on) prof_record(cls, false, s);
    }
};

}  // namespace wust
