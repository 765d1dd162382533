// Block-to-work mapping aware of the MI355X's 8 XCDs.
#pragma once

#include <hip/hip_runtime.h>

namespace uaamg {
namespace dev {

// Workgroups are dispatched to the 8 XCDs round-robin by blockIdx.x, and each
// XCD has its own L2. The logical block index gives XCD x a contiguous range of
// the grid (a bijection for any grid size: the first gridDim % 8 XCDs take one
// block more), so neighbouring leaves in tree order - the halo rows a stencil
// re-reads - are mostly processed by the same XCD and hit its L2.
__device__ __forceinline__ int logical_block() {
    const int b = int(blockIdx.x), nb = int(gridDim.x);
    const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

}  // namespace dev
}  // namespace uaamg
