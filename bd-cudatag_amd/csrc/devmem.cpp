// Device-buffer helpers for callers of the device-pointer entry points
// (tests and benchmarks that stage frames in HBM).  Built as its own small
// library against the same HIP runtime as libbdcudatag.so, so pointers it
// returns are valid for bd_detector_detect_device / bd_batch_detect_device /
// bd_server_submit in the same process.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#define BDM_API extern "C" __attribute__((visibility("default")))

BDM_API void* bdm_alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes == 0 ? 1 : bytes) != hipSuccess) return nullptr;
    return p;
}

BDM_API int32_t bdm_free(void* p) { return hipFree(p) == hipSuccess ? 0 : -1; }

BDM_API int32_t bdm_upload(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}

BDM_API int32_t bdm_download(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

BDM_API int32_t bdm_copy(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) == hipSuccess ? 0 : -1;
}

BDM_API int32_t bdm_sync(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : -1; }

BDM_API int32_t bdm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
