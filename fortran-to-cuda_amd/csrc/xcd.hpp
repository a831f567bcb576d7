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

// The tile of this workgroup in a launch of one workgroup per tile, tiles numbered x fastest, then y, then
// z: in launch order (order 0) or XCD-contiguous (order 1).
__device__ __forceinline__ unsigned launch_tile(int order)
{
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned t = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    return order ? xcd_contiguous(t, gx * gy * gridDim.z) : t;
}

// The tile of this workgroup in a one-call contraction launch over grid (mtiles, ntiles, batch). Consecutive
// tiles in launch order would land in eight different L2s; the XCD-contiguous order gives each XCD one
// contiguous run of the tile order (mt fastest), so the tiles that re-read one operand panel (same nt and
// b, or same mt and b) run under one L2.
__device__ __forceinline__ void gemm_block_tile(int &mt, int &nt, long long &b)
{
    const unsigned mtiles = gridDim.x, ntiles = gridDim.y;
    const unsigned L = launch_tile(1);
    mt = (int)(L % mtiles);
    nt = (int)((L / mtiles) % ntiles);
    b = (long long)(L / (mtiles * ntiles));
}

// A persistent banded walk of `total` tiles by `workgroups` workgroups (a positive multiple of eight): XCD x
// owns the contiguous range [total x / 8, total (x+1) / 8), and its workgroups take consecutive slots of
// that range in step, so the tiles resident on one XCD at a time form a band. Workgroup wg takes the tiles
// first, first + step, ... below end; every tile belongs to one range and one slot of it.
struct TileWalk {
    unsigned first, end, step;
};

__device__ __forceinline__ TileWalk banded_walk(unsigned total, unsigned workgroups, unsigned wg)
{
    const unsigned x = wg % kXcds, slot = wg / kXcds;
    const unsigned lo = (unsigned)((unsigned long long)total * x / kXcds);
    const unsigned hi = (unsigned)((unsigned long long)total * (x + 1) / kXcds);
    return TileWalk{lo + slot, hi, workgroups / kXcds};
}

}  // namespace ftc
