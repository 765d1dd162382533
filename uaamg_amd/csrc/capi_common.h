// Error reporting shared by the C ABI translation units.
#pragma once

#include <exception>
#include <new>
#include <string>

#include "../../include/uaamg_capi.h"
#include "leaf_topology.h"
#include "poisson_solver.h"
#include "viscosity_solver.h"

namespace uaamg {
namespace capi {
// sets the calling thread's last error and returns code
int fail(int code, const std::string& msg);
// validated tree-ordered copy of a caller grid with ncomp components
bool to_host_grid(const uaamg_grid* g, int ncomp, HostGrid& out, const char* what);
// validated options and the host hierarchy of a problem (shared by the
// create entry points and the server); 0 or a uaamg status with the error set
// device_resident: on a device handle (and without flags bit 8) the whole
// hierarchy is built on the device and its coefficient bricks stay there
// (hierarchy_device.h); otherwise the host holds them.
int prepare_poisson(const uaamg_grid* liquid_phi, const uaamg_grid* face_weights, float dt, float dx,
                    const uaamg_options* opts, PoissonHierarchyHost& hier, SolverOptions& so,
                    bool device_resident = false);
int prepare_viscosity(const uaamg_grid* viscosity, const uaamg_grid* liquid_sdf, const uaamg_grid* solid_sdf,
                      float dt, float rho, float dx, const uaamg_viscosity_options* opts, ViscHierarchyHost& hh,
                      ViscOptions& vo);
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
