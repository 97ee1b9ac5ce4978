// Internals shared by the translation units of libdqcrt.so (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include "dogqc_amd.h"

namespace dqcrt {

// Record a message for dqc_last_error() on this thread and return `code`.
int fail(int code, const char* fmt, ...);
int hipFail(hipError_t e, const char* what);

// A loaded plan module: its three C-ABI entry points.
struct Module {
    void* handle;
    const char* (*describe)(void);
    int (*layout)(const dqc_plan_inputs*, const dqc_plan_tuning*, dqc_plan_layout*);
    int (*enqueue)(const dqc_plan_inputs*, const dqc_plan_tuning*, void*, uint64_t, void*, void* const*);
};

// After a run: enlarge in `tun` every table the status reports too small
// (hash tables grow fourfold, since the number of keys that did not fit is
// unknown; multimaps, temp tables, long-list queues and the result get
// their exact counts).  Returns true when the run has to be repeated.
bool adjustForRetry(const dqc_plan_layout& L, const dqc_plan_status& st, dqc_plan_tuning* tun);

}  // namespace dqcrt

#define HIP_CHECK(expr, what)                                      \
    do {                                                           \
        hipError_t _e = (expr);                                    \
        if (_e != hipSuccess) return dqcrt::hipFail(_e, what);     \
    } while (0)
