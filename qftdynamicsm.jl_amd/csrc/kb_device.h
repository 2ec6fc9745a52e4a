// kb_device.h -- device-side layout shared by the kernels of libqftkb.
//
// Two-time history layout in HBM ("mode-block" layout)
//   The reference stores F and r as a lower triangle of N = NstepsinMemory
//   ring slots, every entry a full lattice (src/DefSimDataElements.jl:
//   169-221).  Here the same slot triangle is kept, but the lattice sites are
//   grouped in blocks of MB = 8 consecutive sites and the block index is the
//   slowest dimension:
//
//       H[mb][pair(slot_i, slot_j)][lane][comp]      (double)
//
//   mb    = site / 8, lane = site % 8,
//   pair  = hi*(hi+1)/2 + lo for slots hi >= lo (packed lower triangle),
//   comp  = 0: F,  1: r (rho in the reference's sign-folded convention:
//           the value stored for slots (hi, lo) is rho(time(hi), time(lo))).
//
//   One pair of one block is 8 sites x (F, r) = 128 bytes, i.e. one L2
//   cache line, and the whole two-time history of one block is a single
//   contiguous stream of N(N+1)/2 lines.  The memory-integral kernel gives
//   one workgroup one block and reads that stream exactly once per step.
//
// Self-energy ring: S[mb][slot][lane][comp], comp 0 = SigF, 1 = Sigr.
#pragma once
#include <hip/hip_runtime.h>

namespace qftkb {

constexpr int MB = 8;  // lattice sites per mode block

__host__ __device__ inline long pair_index(int a, int b) {
  const int hi = a > b ? a : b;
  const int lo = a > b ? b : a;
  return (long)hi * (hi + 1) / 2 + lo;
}

// double offset of (block, pair, lane), component F; component r is +1
__host__ __device__ inline size_t hoff(int mb, long pair, int lane, long NP) {
  return ((((size_t)mb * (size_t)NP + (size_t)pair) * MB) + lane) * 2;
}

__host__ __device__ inline size_t soff(int mb, int slot, int lane, int N) {
  return ((((size_t)mb * (size_t)N) + slot) * MB + lane) * 2;
}

}  // namespace qftkb
