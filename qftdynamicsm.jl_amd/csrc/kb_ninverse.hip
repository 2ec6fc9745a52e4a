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
                      int ibeg, int iend, int* next_item, int N, int Rmax, int smin, int L, int excl, int lane,
                      int worker, int nworkers) {
  const int tid = threadIdx.x;
  int item = ibeg + worker;
  while (item < iend) {
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

// One site block's view of the step: slots, coefficients and the rows of
// its triangles and ring for this lane.
struct NinvBlock {
  int N, Rmax, smin, L, sT, lane, worker, nworkers;
  double dt, a3, a6;
  const double* PFb;
  const double* PRb;
  double* IRb;
  double* Ib;
};

__device__ __forceinline__ NinvBlock ninv_block(const NinvArgs& a, const StepState& ss, int mb) {
  NinvBlock b;
  b.N = a.N;
  b.Rmax = ss.Rmax;
  b.smin = ss.s_min;
  b.L = ss.L;
  b.sT = ss.s_tm1;  // slot of T = t - 1
  b.lane = threadIdx.x & (MB - 1);
  b.worker = threadIdx.x >> 3;
  b.nworkers = blockDim.x >> 3;
  b.dt = a.dt;
  b.a3 = a.lam / 3.0;
  b.a6 = a.lam / 6.0;
  b.PFb = a.PF + poff(mb, 0, b.lane, a.PNP);
  b.PRb = a.PR + poff(mb, 0, b.lane, a.PNP);
  b.IRb = a.IR + poff(mb, 0, b.lane, a.PNP);
  b.Ib = a.Iring + soff(mb, 0, b.lane, a.N);
  return b;
}

// before sweep 1: cv = Fr(T, x) (physical sign) for live x != T, acc = 0;
// row T of IR is not written yet and takes no part
__device__ __forceinline__ void phase_fr(const NinvBlock& b, double* cv, double* acc) {
  for (int x = b.worker; x < b.N; x += b.nworkers) {
    const int rk = rank_of(x, b.smin, b.N);
    double v = 0.0;
    if (x < b.Rmax && rk < b.L && x != b.sT) {
      const double pr = b.PRb[prel(ppair_off(b.sT, x, b.N))];
      v = b.sT >= x ? pr : -pr;
    }
    cv[x * MB + b.lane] = v;
    acc[x * MB + b.lane] = 0.0;
  }
}

// after sweep 1: Ir(T, t') from the column sums -- row T of IR, the ring's
// Ir -- and cv = w Ir for sweep 2; acc = 0
__device__ __forceinline__ void phase_ir(const NinvBlock& b, double* cv, double* acc) {
  for (int x = b.worker; x < b.N; x += b.nworkers) {
    const int rk = rank_of(x, b.smin, b.N);
    double w_ir = 0.0;
    if (x < b.Rmax && rk < b.L) {
      double ir = 0.0;
      if (x != b.sT) {
        const long p = ppair_off(b.sT, x, b.N);
        const double pr = b.PRb[prel(p)];
        ir = b.a3 * ((b.sT >= x ? pr : -pr) - b.dt * acc[x * MB + b.lane]);
        b.IRb[prel(p)] = ir;
      }
      b.Ib[(size_t)x * MB * 2 + 1] = ir;
      w_ir = (rk == 0 ? 0.5 : 1.0) * ir;
    }
    cv[x * MB + b.lane] = w_ir;
    acc[x * MB + b.lane] = 0.0;
  }
}

// after sweep 2: d = lambda/6 [F2kr2(T, t') - dt sum_z w_z Ir(T,z) F2kr2(z,t')]
// into cv; acc = 0
__device__ __forceinline__ void phase_d(const NinvBlock& b, double* cv, double* acc) {
  for (int x = b.worker; x < b.N; x += b.nworkers) {
    const int rk = rank_of(x, b.smin, b.N);
    double d = 0.0;
    if (x < b.Rmax && rk < b.L) d = b.a6 * (b.PFb[prel(ppair_off(b.sT, x, b.N))] - b.dt * acc[x * MB + b.lane]);
    cv[x * MB + b.lane] = d;
    acc[x * MB + b.lane] = 0.0;
  }
}

// d'(t') = d(t') + c/2 d(tmin) Fr(tmin, t') for t' > tmin; 0 at tmin
__device__ __forceinline__ void phase_dprime(const NinvBlock& b, double* cv, double dmin) {
  const double half_c = 0.5 * b.a3 * b.dt;
  for (int x = b.worker; x < b.N; x += b.nworkers) {
    const int rk = rank_of(x, b.smin, b.N);
    double dp = 0.0;
    if (x < b.Rmax && rk < b.L && rk > 0) {
      const double pr = b.PRb[prel(ppair_off(b.smin, x, b.N))];
      dp = cv[x * MB + b.lane] + half_c * dmin * (b.smin >= x ? pr : -pr);
    }
    cv[x * MB + b.lane] = dp;
  }
}

// after sweep 3: IF(T, t') = d'(t') - dt sum_{tmin<z<t'} d'(z) Ir(t', z)
// into the ring; IF(T, tmin) = d(tmin) (written by the caller if skip_min)
__device__ __forceinline__ void phase_if(const NinvBlock& b, const double* cv, const double* acc, double dmin,
                                         bool skip_min) {
  for (int x = b.worker; x < b.N; x += b.nworkers) {
    const int rk = rank_of(x, b.smin, b.N);
    if (x < b.Rmax && rk < b.L && !(skip_min && rk == 0))
      b.Ib[(size_t)x * MB * 2] = rk == 0 ? dmin : cv[x * MB + b.lane] - b.dt * acc[x * MB + b.lane];
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
  const NinvBlock b = ninv_block(a, ss, blockIdx.x);
  const int tid = threadIdx.x;

  // sweep 1: column sums of IR with Fr(T, i)
  phase_fr(b, cv, acc);
  if (tid == 0) next_item = b.nworkers;
  __syncthreads();
  sweep<SW_COL>(b.IRb, cv, acc, a.items, 0, a.nitems, &next_item, N, b.Rmax, b.smin, b.L, b.sT, b.lane, b.worker,
                b.nworkers);
  __syncthreads();
  phase_ir(b, cv, acc);
  if (a.passes == NINV_PASS_IR) return;
  if (tid == 0) next_item = b.nworkers;
  __syncthreads();
  // sweep 2: symmetric product with F2kr2
  sweep<SW_SYM>(b.PFb, cv, acc, a.items, 0, a.nitems, &next_item, N, b.Rmax, b.smin, b.L, -1, b.lane, b.worker,
                b.nworkers);
  __syncthreads();
  phase_d(b, cv, acc);
  __syncthreads();
  const double dmin = cv[b.smin * MB + b.lane];  // d(tmin) = IF(T, tmin)
  if (tid == 0) next_item = b.nworkers;
  __syncthreads();
  phase_dprime(b, cv, dmin);
  __syncthreads();
  // sweep 3: row sums of IR with d'
  sweep<SW_ROW>(b.IRb, cv, acc, a.items, 0, a.nitems, &next_item, N, b.Rmax, b.smin, b.L, -1, b.lane, b.worker,
                b.nworkers);
  __syncthreads();
  phase_if(b, cv, acc, dmin, false);
}

// Split form for windows with few site blocks (cubic-orbit storage): each
// sweep's work items of a site block are cut into nsplit ranges, one
// workgroup each, which write their accumulators to part; the finish kernel
// of the sweep sums them and runs the phase that follows, passing the next
// coefficient vector through cvg [site block][slot][lane].
template <int SWEEP>
__global__ __launch_bounds__(NINV_THREADS) void ninv_split_sweep(NinvArgs a, int nsplit, double* __restrict__ part,
                                                                 const double* __restrict__ cvg) {
  extern __shared__ double smem[];
  const StepState ss = *a.st;
  const int N = a.N;
  double* acc = smem;
  double* cv = acc + N * MB;
  __shared__ int next_item;
  const int mb = blockIdx.x / nsplit;
  const int sp = blockIdx.x - mb * nsplit;
  const int ibeg = (int)((long)a.nitems * sp / nsplit);
  const int iend = (int)((long)a.nitems * (sp + 1) / nsplit);
  const NinvBlock b = ninv_block(a, ss, mb);
  const int tid = threadIdx.x;
  if (SWEEP == 1) {
    phase_fr(b, cv, acc);
  } else {
    for (int x = b.worker; x < N; x += b.nworkers) {
      cv[x * MB + b.lane] = cvg[((size_t)mb * N + x) * MB + b.lane];
      acc[x * MB + b.lane] = 0.0;
    }
  }
  if (tid == 0) next_item = ibeg + b.nworkers;
  __syncthreads();
  if (SWEEP == 1)
    sweep<SW_COL>(b.IRb, cv, acc, a.items, ibeg, iend, &next_item, N, b.Rmax, b.smin, b.L, b.sT, b.lane, b.worker,
                  b.nworkers);
  else if (SWEEP == 2)
    sweep<SW_SYM>(b.PFb, cv, acc, a.items, ibeg, iend, &next_item, N, b.Rmax, b.smin, b.L, -1, b.lane, b.worker,
                  b.nworkers);
  else
    sweep<SW_ROW>(b.IRb, cv, acc, a.items, ibeg, iend, &next_item, N, b.Rmax, b.smin, b.L, -1, b.lane, b.worker,
                  b.nworkers);
  __syncthreads();
  double* out = part + (size_t)blockIdx.x * N * MB;
  for (int i = tid; i < N * MB; i += blockDim.x) out[i] = acc[i];
}

template <int SWEEP>
__global__ __launch_bounds__(NINV_FINISH_THREADS) void ninv_split_finish(NinvArgs a, int nsplit,
                                                                         const double* __restrict__ part,
                                                                         double* __restrict__ cvg) {
  extern __shared__ double smem[];
  const StepState ss = *a.st;
  const int N = a.N;
  double* acc = smem;
  double* cv = acc + N * MB;
  const int mb = blockIdx.x;
  const NinvBlock b = ninv_block(a, ss, mb);
  const double* in = part + (size_t)mb * nsplit * N * MB;
  for (int i = threadIdx.x; i < N * MB; i += blockDim.x) {
    double v = 0.0;
    for (int k = 0; k < nsplit; ++k) v += in[(size_t)k * N * MB + i];
    acc[i] = v;
  }
  if (SWEEP == 3) {
    for (int i = threadIdx.x; i < N * MB; i += blockDim.x) cv[i] = cvg[(size_t)mb * N * MB + i];
  }
  __syncthreads();
  if (SWEEP == 1) {
    phase_ir(b, cv, acc);
  } else if (SWEEP == 2) {
    phase_d(b, cv, acc);
    __syncthreads();
    const double dmin = cv[b.smin * MB + b.lane];
    __syncthreads();
    phase_dprime(b, cv, dmin);
    // IF(T, tmin) = d(tmin), before d' replaces d
    if (b.worker == 0 && b.smin < b.Rmax && b.L > 0) b.Ib[(size_t)b.smin * MB * 2] = dmin;
  } else {
    phase_if(b, cv, acc, 0.0, true);
    return;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < N * MB; i += blockDim.x) cvg[(size_t)mb * N * MB + i] = cv[i];
}

template __global__ void ninv_split_sweep<1>(NinvArgs, int, double*, const double*);
template __global__ void ninv_split_sweep<2>(NinvArgs, int, double*, const double*);
template __global__ void ninv_split_sweep<3>(NinvArgs, int, double*, const double*);
template __global__ void ninv_split_finish<1>(NinvArgs, int, const double*, double*);
template __global__ void ninv_split_finish<2>(NinvArgs, int, const double*, double*);
template __global__ void ninv_split_finish<3>(NinvArgs, int, const double*, double*);

}  // namespace qftkb
