// Shared host/device helpers of libck_hip.so: the thread-local error message
// behind ck_last_error(), argument checks that return a ck_status, and the
// launch check.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "ck_tail.h"

namespace ck {

// Sets the calling thread's last error message and returns `status`.
int32_t fail(int32_t status, const char* fmt, ...);

// Clears the last error (successful calls leave the previous message alone,
// as errno does; callers read it only after a nonzero status).
void clear_error();

// Maps a HIP error to CK_EHIP with the runtime's message; CK_OK otherwise.
int32_t hip_status(hipError_t err, const char* what);

// Checks the launch that was just enqueued.
inline int32_t launched(const char* what) {
    return hip_status(hipGetLastError(), what);
}

// Device attributes of the current device, queried once per device.
struct DeviceInfo {
    int device;
    int num_cus;
    size_t lds_per_block;
};
int32_t device_info(DeviceInfo* out);

constexpr float kInf = __builtin_huge_valf();

// Validation of an LDPC code / method (ck_ldpc.hip); CK_OK or CK_EINVAL.
int32_t check_code(const ck_ldpc_code* c);
int32_t check_method(const ck_ldpc_method* m);

}  // namespace ck

#define CK_REQUIRE(cond, ...)                          \
    do {                                               \
        if (!(cond)) return ck::fail(CK_EINVAL, __VA_ARGS__); \
    } while (0)

#define CK_HIP(expr, what)                                       \
    do {                                                         \
        int32_t ck_st_ = ck::hip_status((expr), (what));         \
        if (ck_st_ != CK_OK) return ck_st_;                      \
    } while (0)
