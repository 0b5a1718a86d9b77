// Validation of caller pointers at the C ABI (host code only).
//
// Every array an entry point hands to a kernel must be device memory of the
// device the work runs on, large enough for what the kernel reads or writes.
// A host pointer (pageable or pinned) or a pointer into a too-small
// allocation is refused with FLF_EINVAL before anything is queued, so that a
// wrong argument becomes an error code instead of a GPU memory fault.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

namespace flf {

// Last error of the entry points that take no context (flf_rl_scale,
// flf_rl_ratio, flf_net_upsample, flf_mfma_gemm_nt, flf_datafit_noalloc);
// flf_last_error(NULL) returns it.  One per host thread.
inline std::string& tls_error() {
    static thread_local std::string e;
    return e;
}

// True if [p, p + bytes) lies in one device allocation on `device`
// (hipMalloc'ed memory, including blocks a caching allocator carved from
// it, or managed memory).  Otherwise false, with the reason in `why`.
inline bool device_span(int device, const void* p, size_t bytes, const char* what, std::string& why) {
    if (!p) {
        why = std::string(what) + ": null pointer";
        return false;
    }
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        why = std::string(what) + ": not device memory (host or unknown pointer)";
        return false;
    }
    if (at.type != hipMemoryTypeDevice && !at.isManaged) {
        why = std::string(what) + ": host memory; device memory is required";
        return false;
    }
    if (at.device != device) {
        why = std::string(what) + ": memory of device " + std::to_string(at.device) + ", work on device " +
              std::to_string(device);
        return false;
    }
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) == hipSuccess && base) {
        const size_t used = (size_t)((const char*)p - (const char*)base);
        if (used > size || size - used < bytes) {
            why = std::string(what) + ": allocation holds " + std::to_string(size - (used > size ? size : used)) +
                  " bytes from the pointer, " + std::to_string(bytes) + " needed";
            return false;
        }
    } else {
        (void)hipGetLastError();
    }
    return true;
}

}  // namespace flf
