// Error state, version and device queries of libck_hip.so.
#include "ck_common.h"

#include <mutex>

namespace ck {

namespace {
thread_local char g_error[512] = "";
}

int32_t fail(int32_t status, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
    return status;
}

void clear_error() { g_error[0] = '\0'; }

int32_t hip_status(hipError_t err, const char* what) {
    if (err == hipSuccess) return CK_OK;
    return fail(CK_EHIP, "%s: %s", what, hipGetErrorString(err));
}

int32_t device_info(DeviceInfo* out) {
    static std::mutex mu;
    static DeviceInfo cache[64];
    static bool have[64] = {};
    int dev = 0;
    CK_HIP(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= 64) return fail(CK_ENOSPACE, "device index %d beyond 63", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (!have[dev]) {
        int cus = 0, lds = 0;
        CK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
               "hipDeviceGetAttribute(CU count)");
        CK_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
               "hipDeviceGetAttribute(LDS per block)");
        cache[dev] = DeviceInfo{dev, cus, static_cast<size_t>(lds)};
        have[dev] = true;
    }
    *out = cache[dev];
    return CK_OK;
}

}  // namespace ck

extern "C" int32_t ck_version(void) { return 10000; }

extern "C" const char* ck_last_error(void) { return ck::g_error; }
