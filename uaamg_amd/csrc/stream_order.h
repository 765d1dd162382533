// Stream ordering at the device entry points.
#pragma once

#include <hip/hip_runtime.h>

namespace uaamg {

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
        if (hipError_t e = hipEventRecord(ev_, nullptr)) return e;
        return hipStreamWaitEvent(s, ev_, 0);
    }

private:
    hipEvent_t ev_ = nullptr;
};

}  // namespace uaamg
