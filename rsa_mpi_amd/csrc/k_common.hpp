// Launch size and the counters the kernels leave for the host after a draw
// step or a refinement.
#pragma once

#include "device_common.hpp"
#include "pools.hpp"

namespace rsa_amd {

constexpr int WG = 256;
static inline unsigned nblocks(int64_t n, int per = WG) { return (unsigned)std::max<int64_t>(1, (n + per - 1) / per); }

// Counters written by kernels and read by the host after a draw step.
enum Ctr : int {
  C_NTOUCHED = 0,
  C_CURSOR,       // partner-list allocation cursor
  C_ACCEPTED,     // Undecided after a sweep (how_many<Real, Undecided>)
  C_CONFLICT,     // Conflict after a sweep
  C_REMAIN,       // Not_Validated after the tag update
  C_OVERFLOW,     // cell capacity exceeded
  C_CHAIN_OVF,    // word window too short for the shots
  C_LABEL_STUCK,  // label passes made no progress
  C_DROPPED,      // centres in ghost cells
  C_IMG_MISMATCH, // periodic image whose computed cell is not the mirrored ghost cell
  C_END_REL,      // word offset after the last shot's coordinates
  C_KEEP,         // uncovered voxels
  C_NCHILD,       // children kept by subdivide
  C_GROUPED,      // voxels binned by cell (k_cover's processing order)
  C_SCATTER_ERR,  // k_scatter found a workgroup's children inconsistent with the counting pass
  C_COMPACT_ERR,  // k_compact found more stored spheres than the packed copy holds
  C_MOVED,        // voxels k_move moved into holes (the compaction's traffic)
  C_OVERLAP,      // end-of-run check: colliding (sphere, neighbour) pairs, each pair seen from both sides
  C_UNCOVERED,    // end-of-run check: sample points a sphere of the smallest radius could still take
  C_L0,           // k_outcomes: predicted, mispredicted, covered, predicted covered, waves, uniform waves
  C_L5 = C_L0 + 5,
  C_NUM
};

}  // namespace rsa_amd
