// kb_device.h -- device-side layout shared by the kernels of libqftkb.
//
// Two-time history layout in HBM ("mode-block, column-block" layout)
//   The reference stores F and r as a lower triangle of N = NstepsinMemory
//   ring slots, every entry a full lattice (src/DefSimDataElements.jl:
//   169-221).  Here the same slot triangle is kept, but
//     * lattice sites are grouped in blocks of MB = 8 consecutive sites and
//       the site block is the slowest dimension, and
//     * inside one site block the triangle is stored in column blocks of
//       CB = 4 slots: column block J holds rows r = CB*J .. N-1 of columns
//       CB*J .. CB*J+3, row by row (the CB*(CB-1)/2 slots above the
//       diagonal of each block are padding):
//
//       H[mb][pair_off(slot_i, slot_j)][lane][comp]      (double)
//
//   mb = site / 8, lane = site % 8, comp 0: F, 1: r (rho in the reference's
//   sign-folded convention: the value stored for slots (hi, lo), hi >= lo,
//   is rho(time(hi), time(lo))).
//
//   One slot of one site block is 8 sites x (F, r) = 128 bytes, one L2
//   cache line.  A 16-row x 4-column tile of the triangle is one
//   contiguous 8 KiB run, and the whole two-time history of a site block is
//   one contiguous region that the memory-integral kernel sweeps in order.
//
// Self-energy ring: S[mb][slot][lane][comp], comp 0 = SigF, 1 = Sigr.
//
// 1/N expansion (NinverseNLO): three more slot triangles with one value per
// pair and site (layout at poff below) -- the bubbles F2kr2 and Fr (Fr
// sign-folded like rho) and the resummed kernel Ir(later, earlier) -- and
// the ring of resummed kernels (IF, Ir)(t-1, t') in the self-energy layout.
#pragma once
#include <hip/hip_runtime.h>

namespace qftkb {

constexpr int MB = 8;  // lattice sites per mode block
constexpr int CB = 4;  // slot columns per storage column block

// first slot of column block J: CB * sum_{J' < J} (N - CB J')
__host__ __device__ inline long colblock_offset(int J, int N) {
  return (long)CB * ((long)N * J - (long)CB * J * (J - 1) / 2);
}

// slots per site block (including the diagonal padding)
__host__ __device__ inline long slot_pairs(int N) { return colblock_offset((N + CB - 1) / CB, N); }

// storage slot of the unordered slot pair {a, b}
__host__ __device__ inline long pair_off(int a, int b, int N) {
  const int hi = a > b ? a : b;
  const int lo = a > b ? b : a;
  const int J = lo / CB;
  return colblock_offset(J, N) + (long)(hi - J * CB) * CB + (lo - J * CB);
}

// position in the window of ring slot x when the window's first time sits in
// slot smin (both in [0, N)): x - smin wrapped once, with no integer division
__host__ __device__ inline int window_pos(int x, int smin, int N) {
  const int rel = x - smin;
  return rel < 0 ? rel + N : rel;
}

// double offset of (block, pair, lane), component F; component r is +1
__host__ __device__ inline size_t hoff(int mb, long pair, int lane, long NP) {
  return ((((size_t)mb * (size_t)NP + (size_t)pair) * MB) + lane) * 2;
}

// One-component slot triangles (the 1/N bubbles and resummed kernel) use
// column blocks of PCB slots; inside one row of a column block, columns
// 2k, 2k+1 of all MB lanes are stored together:
//   P[mb][pp / PCB][(pp % PCB) / 2][lane][pp % 2],  pp = ppair_off(slot_i, slot_j),
// so one 16-byte load per lane reads two columns and the MB lanes of a
// worker read 128 contiguous bytes.
constexpr int PCB = 4;

__host__ __device__ inline long pcolblock_offset(int J, int N) {
  return (long)PCB * ((long)N * J - (long)PCB * J * (J - 1) / 2);
}
__host__ __device__ inline long pslot_pairs(int N) { return pcolblock_offset((N + PCB - 1) / PCB, N); }
__host__ __device__ inline long ppair_off(int a, int b, int N) {
  const int hi = a > b ? a : b;
  const int lo = a > b ? b : a;
  const int J = lo / PCB;
  return pcolblock_offset(J, N) + (long)(hi - J * PCB) * PCB + (lo - J * PCB);
}
// offset of pair pp relative to poff(mb, 0, lane, PNP)
__host__ __device__ inline size_t prel(long pp) {
  return (size_t)(pp & ~(long)(PCB - 1)) * MB + (size_t)((pp & (PCB - 1)) >> 1) * (2 * MB) + (size_t)(pp & 1);
}
__host__ __device__ inline size_t poff(int mb, long pp, int lane, long PNP) {
  return (size_t)mb * (size_t)PNP * MB + (size_t)lane * 2 + prel(pp);
}

__host__ __device__ inline size_t soff(int mb, int slot, int lane, int N) {
  return ((((size_t)mb * (size_t)N) + slot) * MB + lane) * 2;
}

}  // namespace qftkb
