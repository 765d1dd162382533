// Stream ordering at the device entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>

namespace uaamg {

// Graph capture and null-stream work exclude each other across the threads
// of a process: HIP rejects work queued on the null stream while a stream of
// the process is capturing ("operation would make the legacy stream depend on
// a capturing blocking stream"), even a non-blocking one, and that invalidates
// the capture. Every graph capture of the library (begin to end) and every
// null-stream operation it issues hold this lock, so concurrent solves on
// several threads (the solve server's workers) never meet. Null-stream work
// that a caller queues from another thread while a handle captures its graphs
// (its first solve, or the first after its smoother changes) is outside the
// lock: such callers set flags bit 1 (no graph replay) or order the two.
inline std::recursive_mutex& capture_lock() {
    static std::recursive_mutex m;
    return m;
}

// A solver running on a non-blocking stream of its own orders that stream
// after the work already queued on the null stream: the default stream of
// callers such as torch, whose kernels may still be producing the device
// vectors handed in. (A caller that supplies the stream orders its own work.)
class CallerOrder {
public:
    CallerOrder() = default;
    CallerOrder(const CallerOrder&) = delete;
    CallerOrder& operator=(const CallerOrder&) = delete;
    ~CallerOrder() {
        if (ev_) (void)hipEventDestroy(ev_);
    }
    hipError_t follow(hipStream_t s) {
        if (!ev_) {
            if (hipError_t e = hipEventCreateWithFlags(&ev_, hipEventDisableTiming)) return e;
        }
        {
            std::lock_guard<std::recursive_mutex> lk(capture_lock());
            if (hipError_t e = hipEventRecord(ev_, nullptr)) return e;
        }
        return hipStreamWaitEvent(s, ev_, 0);
    }

private:
    hipEvent_t ev_ = nullptr;
};

}  // namespace uaamg
