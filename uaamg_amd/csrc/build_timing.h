// Optional wall-clock timing of handle construction stages. With
// UAAMG_BUILD_TIMING=1 in the environment each lap() prints
// "[uaamg build] <stage>: <ms>" to stderr; otherwise it does nothing.
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>

namespace uaamg {

class StageClock {
public:
    StageClock() : on_(enabled()), t_(std::chrono::steady_clock::now()) {}
    void lap(const char* stage) {
        if (!on_) return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[uaamg build] %s: %.2f ms\n", stage,
                     std::chrono::duration<double, std::milli>(now - t_).count());
        t_ = now;
    }

private:
    static bool enabled() {
        static const bool on = [] {
            const char* e = std::getenv("UAAMG_BUILD_TIMING");
            return e && e[0] == '1';
        }();
        return on;
    }
    bool on_;
    std::chrono::steady_clock::time_point t_;
};

}  // namespace uaamg
