// timer.hpp -- optional per-kernel HIP-event timing (mwgemm_timer_* in the C ABI).
#pragma once
#include <hip/hip_runtime.h>

namespace mw {

enum KernelClass {
    KC_PLAIN_REAL = 0,
    KC_PLAIN_CPLX,
    KC_BETA_ONLY,
    KC_OZ_EXPONENT,
    KC_OZ_SPLIT,
    KC_OZ_GEMM,
    KC_OZ_CRT,
    KC_BATCH_PLAIN,
    KC_BATCH_OZ_SPLIT,
    KC_BATCH_OZ_GEMM,
    KC_BATCH_OZ_CRT,
    KC_GEMMTR,
    KC_GEMV,
    KC_COUNT
};

// Returns a slot id (or -1 when timing is off); records the start event on s.
int timer_begin(int cls, hipStream_t s);
// Records the stop event of slot id on s (no-op for id < 0).
void timer_end(int id, hipStream_t s);

// While alive, the kernel timers record nothing on the calling thread (their events are
// created per launch, which the no-allocation entry point must not do).
struct TimerPause {
    bool prev;
    TimerPause();
    ~TimerPause();
};

struct ScopedTimer {
    int id;
    hipStream_t s;
    ScopedTimer(int cls, hipStream_t st) : id(timer_begin(cls, st)), s(st) {}
    ~ScopedTimer() { timer_end(id, s); }
};

}  // namespace mw
