// Error reporting shared by the C entry points: every call returns 0 or a negative code and
// leaves a thread-local message for aldvc_last_error().
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#define ALDVC_ABI_VERSION 2

namespace aldvc {

inline std::string& err_slot() {
    static thread_local std::string s;
    return s;
}

inline void clear_error() { err_slot().clear(); }

inline int fail(int code, const char* msg) {
    err_slot() = msg;
    return code;
}

inline const char* last_error_cstr() { return err_slot().c_str(); }

// Status of the launches issued so far on this thread (does not synchronise).
inline int launch_status() {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        err_slot() = std::string("HIP launch failed: ") + hipGetErrorString(e);
        return -5;
    }
    return 0;
}

// Optional timing of the main kernel of each solver call on this thread: when enabled, events
// are recorded around that launch on the call's stream, and kernel_ms() waits for the second one.
struct KernelTimer {
    bool on = false, have = false;
    hipEvent_t a = nullptr, b = nullptr;
    // per-workgroup start / end wall-clock ticks of the next fused 12-DOF launches (diagnostics)
    unsigned long long* wg_clock = nullptr;
    int64_t wg_capacity = 0;  // workgroups the buffer holds
};

inline KernelTimer& timer() {
    static thread_local KernelTimer t;
    return t;
}

inline void timer_begin(hipStream_t st) {
    KernelTimer& t = timer();
    if (t.on) (void)hipEventRecord(t.a, st);
}

inline void timer_end(hipStream_t st) {
    KernelTimer& t = timer();
    if (t.on) t.have = hipEventRecord(t.b, st) == hipSuccess;
}

}  // namespace aldvc
