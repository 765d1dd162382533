// Error reporting shared by the C ABI translation units.
#pragma once

#include <exception>
#include <new>
#include <string>

#include "../../include/uaamg_capi.h"
#include "leaf_topology.h"

namespace uaamg {
namespace capi {
// sets the calling thread's last error and returns code
int fail(int code, const std::string& msg);
// validated tree-ordered copy of a caller grid with ncomp components
bool to_host_grid(const uaamg_grid* g, int ncomp, HostGrid& out, const char* what);
}  // namespace capi
}  // namespace uaamg

#define UAAMG_GUARD_BEGIN try {
#define UAAMG_GUARD_END                                                        \
    }                                                                          \
    catch (const std::bad_alloc&) {                                            \
        return uaamg::capi::fail(UAAMG_ENOMEM, "host allocation failed");      \
    }                                                                          \
    catch (const std::exception& e) {                                          \
        return uaamg::capi::fail(UAAMG_EINVAL, e.what());                      \
    }                                                                          \
    catch (...) {                                                              \
        return uaamg::capi::fail(UAAMG_EINVAL, "unexpected error");            \
    }
