// Host-side error plumbing shared by the library's translation units: every
// C entry point runs its body under guarded(), which turns a thrown LjxFail
// into an LJX_ERR_* status and a thread-local message (no C++ exception
// crosses the C ABI).
#pragma once

#include "../../include/lajolla_amd.h"

#include <hip/hip_runtime.h>

#include <exception>
#include <new>
#include <string>

namespace ljx {

inline thread_local std::string g_last_error;

struct LjxFail {
    int code;
    std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw LjxFail{code, msg}; }

template <class F>
int guarded(F f) {
    try {
        f();
        return LJX_OK;
    } catch (const LjxFail &e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return LJX_ERR_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return LJX_ERR_INVALID_ARGUMENT;
    }
}

}  // namespace ljx

#define HIPX(call)                                                                                    \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw ::ljx::LjxFail{e_ == hipErrorOutOfMemory ? LJX_ERR_OUT_OF_MEMORY : LJX_ERR_HIP,     \
                                 std::string(#call) + ": " + hipGetErrorString(e_)};                  \
    } while (0)
