// Per-kernel-class HIP-event timing (tooling; off unless enabled).
#pragma once
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
    PC_COUNT
};

bool prof_on(int cls);
void prof_record(int cls, bool begin, hipStream_t s);
void prof_flush();

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
