// kb_ninverse.hip -- resummed bubble chains of the 1/N expansion at NLO.
//
// The reference computes, at every step and for every momentum p, the row
// tmone = t-1 of two resummed kernels (calcIrs!, calcIFs!: TwoPIScalar_
// Evolution.jl:170-206 and 354-436; in momentum space as in the CPUred2
// back-end, :252-292 and :539-597).  With T = t-1, a = lambda/3 and the
// bubbles Fr(z,t') (antisymmetric) and F2kr2(z,t') (symmetric):
//
//   Ir(T,t') = a [Fr(T,t') - dt sum_{z=t'+1}^{T-1} Ir(T,z) Fr(z,t')]
//   IF(T,t') = lambda/6 [F2kr2(T,t') - dt sum_{z=tmin}^{T-1} w_z Ir(T,z) F2kr2(z,t')]
//            + lambda/3 dt sum_{z=tmin}^{t'-1} w_z IF(T,z) Fr(z,t')
//
// (w = 1/2 at tmin, 1 otherwise).  Both are triangular solves along the
// window, sequential in t' (downwards for Ir, upwards for IF).  Here they
// become three bandwidth-bound sweeps with no sequential chain:
//
//  * Ir is the resolvent of the strictly lower-triangular kernel K = Fr:
//    R = a K (1 + a dt K)^-1 = (1 + a dt K)^-1 a K.  The second form gives
//    the same row from the rows already computed,
//      Ir(T,t') = a [Fr(T,t') - dt sum_{z=t'+1}^{T-1} Fr(T,z) Ir(z,t')],
//    and Ir(z,t') depends only on times in [t', z], not on the window's
//    start, so every row computed at an earlier step stays valid.  Rows are
//    kept in a slot triangle IR(later, earlier); sweep 1 forms the column
//    sums sum_i Fr(T,i) IR(i,j).
//  * The IF system (1 - c W K')y = d, c = a dt, K'(z,t') = Fr(z,t') for
//    z < t', has the inverse G = 1 - dt R^T over the window when W = 1.  The
//    weight 1/2 at tmin only touches row tmin; since y(tmin) = d(tmin), it
//    moves into the source: d'(t') = d(t') + c/2 d(tmin) Fr(tmin,t'), and
//      y(t') = d'(t') - dt sum_{tmin<z<t'} d'(z) IR(t',z)   (t' > tmin).
//    Sweep 2 is the symmetric product with F2kr2 that gives d, sweep 3 the
//    row sums with IR.
//
// Each sweep reads one slot triangle once (8 bytes per pair and site), with
// the work items and the two-batch streaming of the memory integral; the
// coefficient vector and the accumulators of one site block sit in LDS.
#include "kb_kernels.h"

namespace qftkb {

namespace {

enum { SW_COL = 0, SW_SYM = 1, SW_ROW = 2 };

__device__ __forceinline__ int rank_of(int s, int smin, int N) {
  const int r = s - smin;
  return r < 0 ? r + N : r;
}

typedef double nv_d2 __attribute__((ext_vector_type(2)));
// the PCB values of one row of a column block (one 16-byte load per column
// pair; p points at the row's first column for this lane), streamed
__device__ __forceinline__ void ld_row(const double* p, double (&v)[PCB]) {
#pragma unroll
  for (int k = 0; k < PCB / 2; ++k) {
    const nv_d2 a = __builtin_nontemporal_load(reinterpret_cast<const nv_d2*>(p + (size_t)k * 2 * MB));
    v[2 * k] = a.x;
    v[2 * k + 1] = a.y;
  }
}

#ifndef KB_NINV_BATCH
#define KB_NINV_BATCH 2
#endif
constexpr int NB_ROWS = KB_NINV_BATCH;  // rows per batch, two batches in flight

// One sweep over the stored pairs (r >= c in slot space) of one triangle of
// this site block.  With i the later and j the earlier time of the pair:
//   SW_COL: acc[j] += cv[i] T      SW_ROW: acc[i] += cv[j] T
//   SW_SYM: acc[r] += cv[c] T and acc[c] += cv[r] T (diagonal once).
// Slots outside the live window, and the slot `excl`, take no part:
// contributions of their pairs are dropped by select, so values not yet
// written for this window never enter a live sum.
template <int MODE>
__device__ __forceinline__ void pair_term(double v, bool ok, bool nrm, double cr, double cc, double& row,
                                          double& col) {
  if (MODE == SW_SYM) {
    row += ok ? cc * v : 0.0;
    col += ok ? cr * v : 0.0;
  } else if (MODE == SW_COL) {
    col += (ok && nrm) ? cr * v : 0.0;
    row += (ok && !nrm) ? cc * v : 0.0;
  } else {
    row += (ok && nrm) ? cc * v : 0.0;
    col += (ok && !nrm) ? cr * v : 0.0;
  }
}

template <int MODE>
__device__ void sweep(const double* __restrict__ Tb, const double* cv, double* acc, const int2* __restrict__ items,
                      int nitems, int* next_item, int N, int Rmax, int smin, int L, int excl, int lane, int worker,
                      int nworkers) {
  const int tid = threadIdx.x;
  int item = worker;
  while (item < nitems) {
    const int2 jk = items[item];
    const int c0 = jk.x * PCB;
    const int r0 = c0 + jk.y * NINV_ROWS;
    const int r1 = min(r0 + NINV_ROWS, Rmax);
    if (c0 < Rmax && r0 < r1) {
      const long cbase = pcolblock_offset(jk.x, N);
      double cc[PCB], col[PCB];
      int rc[PCB];
#pragma unroll
      for (int jc = 0; jc < PCB; ++jc) {
        const int c = c0 + jc;
        rc[jc] = c == excl ? N : rank_of(c, smin, N);
        const bool ok = c < Rmax && rc[jc] < L;
        cc[jc] = ok ? cv[min(c, N - 1) * MB + lane] : 0.0;
        if (!ok) rc[jc] = N;  // never live
        col[jc] = 0.0;
      }
      int r = r0;
      // rows that cross the diagonal of the column block
      for (; r < r1 && r < c0 + PCB; ++r) {
        const int rr = r == excl ? N : rank_of(r, smin, N);
        const bool okr = rr < L;
        const double cr = okr ? cv[r * MB + lane] : 0.0;
        double row = 0.0;
        const long pb = cbase + (long)(r - c0) * PCB;
        for (int jc = 0; jc < PCB; ++jc) {
          const int c = c0 + jc;
          if (c > r) break;
          const double v = Tb[prel(pb + jc)];
          const bool ok = okr && rc[jc] < L;
          if (c == r) {
            if (MODE == SW_SYM) row += ok ? cr * v : 0.0;
            continue;
          }
          pair_term<MODE>(v, ok, rr > rc[jc], cr, cc[jc], row, col[jc]);
        }
        atomicAdd(&acc[r * MB + lane], row);
      }
      // full rows, NB_ROWS per batch, two batches in flight; rows past r1 are
      // clamped to r1 - 1 (re-read, not used)
      if (r < r1) {
        double X[NB_ROWS][PCB], Y[NB_ROWS][PCB];
        auto load = [&](double (&b)[NB_ROWS][PCB], int rs) {
#pragma unroll
          for (int h = 0; h < NB_ROWS; ++h) {
            const int rr = min(rs + h, r1 - 1);
            ld_row(Tb + prel(cbase + (long)(rr - c0) * PCB), b[h]);  // row start: a multiple of PCB
          }
        };
        auto process = [&](const double (&b)[NB_ROWS][PCB], int rs) {
#pragma unroll
          for (int h = 0; h < NB_ROWS; ++h) {
            const int rw = rs + h;
            if (rw < r1) {
              const int rr = rw == excl ? N : rank_of(rw, smin, N);
              const bool okr = rr < L;
              const double cr = okr ? cv[rw * MB + lane] : 0.0;
              double row = 0.0;
#pragma unroll
              for (int jc = 0; jc < PCB; ++jc)
                pair_term<MODE>(b[h][jc], okr && rc[jc] < L, rr > rc[jc], cr, cc[jc], row, col[jc]);
              atomicAdd(&acc[rw * MB + lane], row);
            }
          }
        };
        load(X, r);
        while (true) {
          load(Y, r + NB_ROWS);
          process(X, r);
          r += NB_ROWS;
          if (r >= r1) break;
          load(X, r + NB_ROWS);
          process(Y, r);
          r += NB_ROWS;
          if (r >= r1) break;
        }
      }
#pragma unroll
      for (int jc = 0; jc < PCB; ++jc) {
        const int c = c0 + jc;
        if (c < Rmax) atomicAdd(&acc[c * MB + lane], col[jc]);
      }
    }
    int nxt = 0;
    if (lane == 0) nxt = atomicAdd(next_item, 1);
    item = __shfl(nxt, tid & ~(MB - 1) & 63);
  }
}

}  // namespace

__global__ __launch_bounds__(NINV_THREADS) void ninv_resum_kernel(NinvArgs a) {
  extern __shared__ double smem[];
  const StepState ss = *a.st;
  const int N = a.N;
  double* acc = smem;          // accumulators per slot, lane
  double* cv = acc + N * MB;   // coefficient vector of the sweep
  __shared__ int next_item;
  const int mb = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & (MB - 1);
  const int worker = tid >> 3;
  const int nworkers = blockDim.x >> 3;
  const int Rmax = ss.Rmax;
  const int smin = ss.s_min;
  const int L = ss.L;
  const int sT = ss.s_tm1;  // slot of T = t - 1
  const double dt = a.dt;
  const double a3 = a.lam / 3.0, a6 = a.lam / 6.0;
  const double* __restrict__ PFb = a.PF + poff(mb, 0, lane, a.PNP);
  const double* __restrict__ PRb = a.PR + poff(mb, 0, lane, a.PNP);
  double* __restrict__ IRb = a.IR + poff(mb, 0, lane, a.PNP);
  double* __restrict__ Ib = a.Iring + soff(mb, 0, lane, N);

  // sweep 1: column sums of IR with Fr(T, i); row T of IR is not written
  // yet and takes no part
  for (int x = worker; x < N; x += nworkers) {
    const int rk = rank_of(x, smin, N);
    double v = 0.0;
    if (x < Rmax && rk < L && x != sT) {
      const double pr = PRb[prel(ppair_off(sT, x, N))];
      v = sT >= x ? pr : -pr;
    }
    cv[x * MB + lane] = v;
    acc[x * MB + lane] = 0.0;
  }
  if (tid == 0) next_item = nworkers;
  __syncthreads();
  sweep<SW_COL>(IRb, cv, acc, a.items, a.nitems, &next_item, N, Rmax, smin, L, sT, lane, worker, nworkers);
  __syncthreads();
  // Ir(T, t'): row T of IR, the ring's Ir, and w Ir for sweep 2
  for (int x = worker; x < N; x += nworkers) {
    const int rk = rank_of(x, smin, N);
    double w_ir = 0.0;
    if (x < Rmax && rk < L) {
      double ir = 0.0;
      if (x != sT) {
        const long p = ppair_off(sT, x, N);
        const double pr = PRb[prel(p)];
        ir = a3 * ((sT >= x ? pr : -pr) - dt * acc[x * MB + lane]);
        IRb[prel(p)] = ir;
      }
      Ib[(size_t)x * MB * 2 + 1] = ir;
      w_ir = (rk == 0 ? 0.5 : 1.0) * ir;
    }
    cv[x * MB + lane] = w_ir;
    acc[x * MB + lane] = 0.0;
  }
  if (a.passes == NINV_PASS_IR) return;
  if (tid == 0) next_item = nworkers;
  __syncthreads();
  // sweep 2: d = lambda/6 [F2kr2(T, t') - dt sum_z w_z Ir(T,z) F2kr2(z,t')]
  sweep<SW_SYM>(PFb, cv, acc, a.items, a.nitems, &next_item, N, Rmax, smin, L, -1, lane, worker, nworkers);
  __syncthreads();
  for (int x = worker; x < N; x += nworkers) {
    const int rk = rank_of(x, smin, N);
    double d = 0.0;
    if (x < Rmax && rk < L) d = a6 * (PFb[prel(ppair_off(sT, x, N))] - dt * acc[x * MB + lane]);
    cv[x * MB + lane] = d;
    acc[x * MB + lane] = 0.0;
  }
  __syncthreads();
  const double dmin = cv[smin * MB + lane];  // d(tmin) = IF(T, tmin)
  if (tid == 0) next_item = nworkers;
  __syncthreads();
  // d'(t') = d(t') + c/2 d(tmin) Fr(tmin, t') for t' > tmin; 0 at tmin
  const double half_c = 0.5 * a3 * dt;
  for (int x = worker; x < N; x += nworkers) {
    const int rk = rank_of(x, smin, N);
    double dp = 0.0;
    if (x < Rmax && rk < L && rk > 0) {
      const double pr = PRb[prel(ppair_off(smin, x, N))];
      dp = cv[x * MB + lane] + half_c * dmin * (smin >= x ? pr : -pr);
    }
    cv[x * MB + lane] = dp;
  }
  __syncthreads();
  // sweep 3: IF(T, t') = d'(t') - dt sum_{tmin<z<t'} d'(z) Ir(t', z)
  sweep<SW_ROW>(IRb, cv, acc, a.items, a.nitems, &next_item, N, Rmax, smin, L, -1, lane, worker, nworkers);
  __syncthreads();
  for (int x = worker; x < N; x += nworkers) {
    const int rk = rank_of(x, smin, N);
    if (x < Rmax && rk < L) Ib[(size_t)x * MB * 2] = rk == 0 ? dmin : cv[x * MB + lane] - dt * acc[x * MB + lane];
  }
}

}  // namespace qftkb
