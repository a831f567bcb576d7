// Workgroup order across the eight XCDs of an MI355X. The dispatcher hands workgroups to the XCDs
// round-robin in launch order: linear workgroup id L runs on XCD L % 8, as that XCD's (L / 8)-th.
// Consecutive ids thus land in eight different L2 caches.
#pragma once
#include <hip/hip_runtime.h>

namespace ftc {

constexpr unsigned kXcds = 8;

// Maps the linear id L of a launch of `total` workgroups to a logical index such that XCD x receives
// the contiguous run [s_x, s_x + c_x), where c_x = total / 8, plus one for the first total % 8 XCDs, and
// s_x is the sum of the earlier counts. The map is a bijection on [0, total), so every logical tile is
// computed exactly once whatever total is.
__device__ __forceinline__ unsigned xcd_contiguous(unsigned L, unsigned total)
{
    const unsigned x = L % kXcds, q = total / kXcds, r = total % kXcds;
    return x * q + (x < r ? x : r) + L / kXcds;
}

}  // namespace ftc
