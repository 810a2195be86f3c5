// Shared definitions of the MI355X EMC engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace emc {

// Orientations handled by one wavefront: one lane per orientation, so a photon's row of
// log W (pixel-major, orientation-contiguous) is one 512-byte coalesced read.
constexpr int kRTile = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;
// MI355X dispatches consecutive workgroups of a launch round-robin over its 8 XCDs, each with
// its own L2.
constexpr int kNumXcd = 8;

// Mixed-precision product tiles (emc_mp.hip): block tile kGemmBM x kGemmBN, kGemmBK halves
// of K per step, 8 waves; kGemmGroupM m-tiles per panel of the block order.
constexpr int kGemmBM = 256;
constexpr int kGemmBN = 256;
constexpr int kGemmBK = 32;
constexpr int kGemmThreads = 512;
constexpr int kGemmGroupM = 8;
// K step of the mixed-precision EMC products (the proposal network's products use kGemmBK)
constexpr int kMpGemmBK = 64;

// Brick merge: bricks of kBrick^3 voxels accumulated in LDS; at most kUnitItems work items
// (orientation, 64-pixel patch) per block.
constexpr int kBrick = 16;
constexpr int kUnitItems = 2048;
// Sparse subsets: pixels per LDS range of the pair tomogram.
constexpr int kPairRange = 64;

// Reference constants (dragonfly/src/max_internal.h:6-8).
constexpr double kProbMin = 1.e-6;
constexpr double kPdiffThresh = 14.;
constexpr double kMaxExpStart = -1.e100;

// A dispatch's grid size is counted in work-items in 32 bits: a 1-D launch of blocks x kBlock
// work-items past 2^32 wraps silently and drops the tail. Kernels whose work can grow past
// that (pairs x 64, items x 64) take a capped grid and stride over their work.
constexpr long kMaxStrideBlocks = 1L << 20;
inline unsigned stride_grid(long blocks) {
    return (unsigned)(blocks < kMaxStrideBlocks ? blocks : kMaxStrideBlocks);
}

}  // namespace emc
