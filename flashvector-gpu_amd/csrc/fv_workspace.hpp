// Per-(device, stream) scratch buffers of the entry points that take no
// workspace argument (fv_workspace.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace fv {

// A buffer of at least the requested size, reserved for the caller until
// release() (or destruction): enqueue every kernel that uses it before then.
class WorkspaceLease {
public:
    WorkspaceLease() = default;
    WorkspaceLease(const WorkspaceLease&) = delete;
    WorkspaceLease& operator=(const WorkspaceLease&) = delete;
    ~WorkspaceLease();
    // Records the end of the caller's work on the stream and frees the lease.
    void release();
    void* ptr = nullptr;

private:
    friend hipError_t stream_workspace(hipStream_t, size_t, WorkspaceLease&);
    void* entry_ = nullptr;
};

// Leases the (current device, stream) buffer, growing it to `bytes` if needed
// (growth waits for the buffer's previous work, then reallocates).
hipError_t stream_workspace(hipStream_t stream, size_t bytes, WorkspaceLease& lease);

}  // namespace fv
