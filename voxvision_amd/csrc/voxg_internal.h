// Services of the main translation unit (voxgpu.hip) that the other HIP
// translation units of libvoxgpu use: the thread's last-error message and
// per-stream device scratch memory.
#pragma once
#include <hip/hip_runtime.h>

#include <stddef.h>

#include <string>

namespace voxg_internal {

// Records msg as the calling thread's voxg_last_error() and returns code.
int fail(int code, const std::string &msg);

// A device buffer of at least `bytes` bytes owned by the (current device,
// stream) pair, grown on demand and kept until the process ends.  Work
// queued on the same stream may use it; work on other streams gets its own.
int scratch(hipStream_t stream, size_t bytes, void **out);

// Frees every buffer the library keeps for (current device, stream); call
// before destroying a stream the library has used.  Waits for the stream.
void release_stream(hipStream_t stream);

}  // namespace voxg_internal

#define VOXG_HIPCHK(expr)                                                  \
    do {                                                                   \
        hipError_t e_ = (expr);                                            \
        if (e_ != hipSuccess)                                              \
            return voxg_internal::fail(                                    \
                VOXG_EIO, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)
