// Host side of the device pressure-gradient update (face_update.h).
#include "face_update.h"

#include <vector>

#include "leaf_topology.h"
#include "poisson_solver.h"

namespace uaamg {

namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

bool build_lookup(int64_t n, const int32_t* org, LeafLookup& lookup, std::string& err, const char* what) {
    std::vector<Coord> o(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        o[size_t(i)] = {org[3 * i], org[3 * i + 1], org[3 * i + 2]};
        if ((o[size_t(i)].x | o[size_t(i)].y | o[size_t(i)].z) & 7) {
            err = std::string(what) + ": leaf origin not a multiple of 8";
            return false;
        }
    }
    lookup.build(o);
    for (size_t i = 0; i < o.size(); ++i)
        if (lookup.find(o[i]) != int32_t(i)) {
            err = std::string(what) + ": duplicate leaf origins";
            return false;
        }
    return true;
}

}  // namespace

int apply_pressure_gradient_device(PoissonDevice& d, const float* pressure_dev, const FaceGridRef* grids,
                                   FaceChannelRef* channels, float dt, float friction, bool values_on_device,
                                   std::string& err) {
    if (!d.on_device()) {
        err = "handle was created without a device (device < 0)";
        return -4;
    }
    hipStream_t s = d.stream();
    if (hipSetDevice(d.device()) != hipSuccess) {
        err = "hipSetDevice failed";
        return -3;
    }
    if (d.follow_caller()) {
        err = d.error();
        return -3;
    }
    auto hip_fail = [&](hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        (void)hipStreamSynchronize(s);
        return -3;
    };
    const PoissonLevelHost& l0 = d.host().levels[0];
    LeafLookup dof_lookup;
    dof_lookup.build(l0.origin);
    // inputs: liquid phi, face weights, solid velocity
    const char* names[3] = {"liquid_phi", "face_weights", "solid_velocity"};
    LeafLookup in_lookup[3];
    DevBuf in_buf[3];
    const float* in_dev[3] = {nullptr, nullptr, nullptr};
    for (int g = 0; g < 3; ++g) {
        const FaceGridRef& r = grids[g];
        if (!build_lookup(r.nleaf, r.origins, in_lookup[g], err, names[g])) return -1;
        if (r.nleaf == 0) continue;
        if (values_on_device) {
            in_dev[g] = r.values;
            continue;
        }
        const size_t bytes = size_t(r.nleaf) * kLeafSize * size_t(r.ncomp) * sizeof(float);
        if (hipError_t e = hipMalloc(&in_buf[g].p, bytes)) return hip_fail(e, "hipMalloc");
        if (hipError_t e = hipMemcpyAsync(in_buf[g].p, r.values, bytes, hipMemcpyHostToDevice, s))
            return hip_fail(e, "upload grid values");
        in_dev[g] = static_cast<const float*>(in_buf[g].p);
    }
    const LevelView& v0 = d.level_view(0);
    for (int c = 0; c < 3; ++c) {
        FaceChannelRef& ch = channels[c];
        LeafLookup vl;
        if (!build_lookup(ch.nleaf, ch.origins, vl, err, "velocity channel")) return -1;
        if (ch.nleaf == 0) continue;
        const size_t n = size_t(ch.nleaf);
        std::vector<int32_t> tab(n * kGradTab);
        for (size_t i = 0; i < n; ++i) {
            const Coord o{ch.origins[3 * i], ch.origins[3 * i + 1], ch.origins[3 * i + 2]};
            Coord lo = o;
            lo[c] -= 8;
            int32_t* t = &tab[i * kGradTab];
            t[kGradPhi] = in_lookup[0].find(o);
            t[kGradPhiLow] = in_lookup[0].find(lo);
            t[kGradDof] = dof_lookup.find(o);
            t[kGradDofLow] = dof_lookup.find(lo);
            t[kGradW] = in_lookup[1].find(o);
            t[kGradSolid] = in_lookup[2].find(o);
        }
        const size_t vbytes = n * kLeafSize * sizeof(float);
        // the leaf tables and the masks in the handle's grow-only buffer (a repeated call allocates
        // nothing; the channels take turns, each synchronised before the next)
        const size_t tab_bytes = (tab.size() * sizeof(int32_t) + 255) & ~size_t(255);
        char* aux = static_cast<char*>(d.aux(tab_bytes + n * 8 * sizeof(uint64_t)));
        if (!aux) return hip_fail(hipErrorOutOfMemory, "device scratch");
        int32_t* tab_d = reinterpret_cast<int32_t*>(aux);
        uint64_t* act_d = reinterpret_cast<uint64_t*>(aux + tab_bytes);
        DevBuf vel_d, upd_d, refl_d;
        if (hipError_t e = hipMemcpyAsync(tab_d, tab.data(), tab.size() * sizeof(int32_t), hipMemcpyHostToDevice, s))
            return hip_fail(e, "upload tables");
        if (hipError_t e = hipMemcpyAsync(act_d, ch.active, n * 8 * sizeof(uint64_t), hipMemcpyHostToDevice, s))
            return hip_fail(e, "upload active masks");
        float* vel = ch.values;
        float* upd = ch.update;
        float* refl = ch.reflected;
        if (!values_on_device) {
            if (hipError_t e = hipMalloc(&vel_d.p, vbytes)) return hip_fail(e, "hipMalloc");
            if (hipError_t e = hipMemcpyAsync(vel_d.p, ch.values, vbytes, hipMemcpyHostToDevice, s))
                return hip_fail(e, "upload velocity");
            vel = static_cast<float*>(vel_d.p);
            if (ch.update) {
                if (hipError_t e = hipMalloc(&upd_d.p, vbytes)) return hip_fail(e, "hipMalloc");
                upd = static_cast<float*>(upd_d.p);
            }
            if (ch.reflected) {
                if (hipError_t e = hipMalloc(&refl_d.p, vbytes)) return hip_fail(e, "hipMalloc");
                refl = static_cast<float*>(refl_d.p);
            }
        }
        GradArgs a{};
        a.tab = tab_d;
        a.phi = in_dev[0];
        a.w = in_dev[1];
        a.sv = in_dev[2];
        a.pressure = pressure_dev;
        a.dmask = v0.mask;
        a.dofbase = v0.dofbase;
        a.phi_bg = grids[0].bg[0];
        for (int k = 0; k < 3; ++k) {
            a.w_bg[k] = grids[1].bg[k];
            a.sv_bg[k] = grids[2].bg[k];
        }
        a.channel = c;
        a.dt = dt;
        a.dx = l0.dx;
        a.friction = friction;
        if (hipError_t e = launch_pressure_gradient(ch.nleaf, a, act_d, vel, upd, refl, s))
            return hip_fail(e, "pressure gradient kernel");
        if (hipError_t e = hipMemcpyAsync(ch.active, act_d, n * 8 * sizeof(uint64_t), hipMemcpyDeviceToHost, s))
            return hip_fail(e, "download active masks");
        if (!values_on_device) {
            if (hipError_t e = hipMemcpyAsync(ch.values, vel, vbytes, hipMemcpyDeviceToHost, s))
                return hip_fail(e, "download velocity");
            if (ch.update)
                if (hipError_t e = hipMemcpyAsync(ch.update, upd, vbytes, hipMemcpyDeviceToHost, s))
                    return hip_fail(e, "download update");
            if (ch.reflected)
                if (hipError_t e = hipMemcpyAsync(ch.reflected, refl, vbytes, hipMemcpyDeviceToHost, s))
                    return hip_fail(e, "download reflected");
        }
        if (hipError_t e = hipStreamSynchronize(s)) return hip_fail(e, "pressure gradient sync");
    }
    return 0;
}

}  // namespace uaamg
