// Device-memory helpers for FFI callers that have no HIP binding of their
// own (ctypes, JNI, cgo): allocate, free, copy, synchronise.
#include "hipibs_internal.hpp"

using namespace hipibs;

extern "C" int hipibs_dev_alloc(int device, uint64_t bytes, void **out)
{
    if (!out)
        return setError(HIPIBS_E_ARG, "null argument");
    *out = nullptr;
    HIPIBS_CHECK(hipSetDevice(device));
    const hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    if (e != hipSuccess)
        return setError(HIPIBS_E_NOMEM, "hipMalloc of " + std::to_string(bytes) +
                                            " bytes failed: " + hipGetErrorString(e));
    return 0;
}

extern "C" int hipibs_dev_free(void *p)
{
    if (p)
        HIPIBS_CHECK(hipFree(p));
    return 0;
}

extern "C" int hipibs_memcpy(void *dst, const void *src, uint64_t bytes, int kind)
{
    if (bytes == 0)
        return 0;
    if (!dst || !src)
        return setError(HIPIBS_E_ARG, "null pointer");
    hipMemcpyKind k;
    switch (kind)
    {
    case 1: k = hipMemcpyHostToDevice; break;
    case 2: k = hipMemcpyDeviceToHost; break;
    case 3: k = hipMemcpyDeviceToDevice; break;
    default: return setError(HIPIBS_E_ARG, "kind must be 1 (H2D), 2 (D2H) or 3 (D2D)");
    }
    HIPIBS_CHECK(hipMemcpy(dst, src, bytes, k));
    return 0;
}

extern "C" int hipibs_dev_memset(void *dst, int value, uint64_t bytes)
{
    if (bytes && !dst)
        return setError(HIPIBS_E_ARG, "null pointer");
    if (bytes)
        HIPIBS_CHECK(hipMemset(dst, value, bytes));
    return 0;
}

extern "C" int hipibs_dev_synchronize(int device)
{
    HIPIBS_CHECK(hipSetDevice(device));
    HIPIBS_CHECK(hipDeviceSynchronize());
    return 0;
}
