// kb_adjoint.hip -- reverse-mode sweep of the 2PI Kadanoff-Baym step.
//
// The reference has no derivative code; this is the adjoint of its evolve!
// (TwoPIScalar_Evolution.jl:924-981) as restated by the forward kernels of
// this library, for the loss
//     L = sum_p wgt(p) n(p, T),
// n measured at T as createMeasurement does (TwoPIScalar_Measurement.jl:
// 38-59), with respect to the coupling lambda (self-energy prefactors and
// the Hartree term) and the three initial rows F(1,1), F(2,1), F(2,2) of
// initSimData! (TwoPIScalar_Initialization.jl:8-29), hence the initial
// occupations n(p, 0).
//
// The adjoint window Hbar has the layout of the history H (rho_bar in the
// same sign-folded convention).  Reverse step t, with the primal window of
// step t resident:
//   1. recompute Sigma(t-1, .) (saving the transforms Fx, rx) and M^2(t);
//   2. adj_row_kernel: adjoint of the diagonal F(t,t) and of the leapfrog
//      row t: R_bar_F, R_bar_r per slot, pushes into rows t-1 and t-2,
//      the Omega adjoint per site block; it first settles the Hartree
//      adjoint of step t+1 into F_bar(t,t);
//   3. adj_reduce_kernel: the self-energy adjoints Sigma_bar of the memory
//      integrals, reduced per slot in LDS from one read of H;
//   4. the self-energy stage run backwards (kb_sigma.hip, ADJ) adds the
//      sunset adjoint into row t-1 of Hbar;
//   5. every AK (8) reverse steps, adj_apply_kernel adds the memory integrals'
//      entry adjoints of those steps to Hbar in one read-modify-write
//      (adj_row_kernel adds the deferred terms of the row it consumes);
//   6. adj_slide_kernel: row t of Hbar is cleared and the column t-N that
//      row t overwrote is put back into H from the tape.
// The tape holds, for every forward step that dropped a column, that
// column's N slot pairs.  With checkpointing, only the window at every C-th
// step is kept and each segment's steps are recomputed (taping) before it
// is reversed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <thread>

#include <unistd.h>

#include "kb_internal.h"

using namespace qftkb;
using namespace qftkb_host;

#pragma GCC visibility push(hidden)
namespace qftkb {

// deferred reverse steps whose memory-integral entry updates are applied
// together: the adjoint window's read-modify-write costs 32 B per entry per
// AK steps, and the row kernel reads the rings of up to AK-1 deferred steps
// per slot of the row it consumes
constexpr int AK = 8;

// Per-step arguments of the reverse kernels.
struct AdjArgs {
  const double* H;
  double* Hbar;
  const double* S;     // Sigma(t-1, .) ring
  double* Rb;          // (R_bar_F, R_bar_r) ring, self-energy layout
  double* Sbar;        // (SigF_bar, Sigr_bar) ring
  const double* kL2;
  const double* m2r;   // [parity][2]: M^2(t), sum_p F(t-1,t-1;p)
  double* P;           // [parity][NB]: Omega adjoint summed over a site block
  double* LP;          // [NB]: running dL/dlambda per site block
  const int2* items;
  int nitems;
  long NP;
  int N, V, NB;
  double dt;
  double hl_ivol;      // lambda (N+2)/(6N) / V: dM^2 / dF
  double h_ivol;       // (N+2)/(6N) / V:        dM^2 / dlambda per unit sum F
  double two_over_lambda;
  int integral;
  int first;           // t == T: no Hartree adjoint pending
  StepState st;        // the forward state of step t
  // deferred memory-integral updates: steps pk[j] with windows starting at
  // ptmin[j], their self-energy and R_bar rings (row kernel: the steps after
  // t not yet applied; apply kernel: the steps being applied)
  int npend;
  int pk[AK], ptmin[AK];
  const double* Sj[AK];
  const double* Rj[AK];
};

namespace {

__device__ __forceinline__ double2 ld2(const double* p) { return *reinterpret_cast<const double2*>(p); }
__device__ __forceinline__ void st2(double* p, double2 v) { *reinterpret_cast<double2*>(p) = v; }

typedef double kb_a2v __attribute__((ext_vector_type(2)));
// streamed once per step: keep them out of the way of the coefficient rings
__device__ __forceinline__ double2 ld2_nt(const double* p) {
  const kb_a2v v = __builtin_nontemporal_load(reinterpret_cast<const kb_a2v*>(p));
  return make_double2(v.x, v.y);
}

// sum of n values by a whole workgroup (fixed order: identical in every
// workgroup that sums the same array)
template <int NT>
__device__ double block_sum(double v, double* red) {
  static_assert(NT % 64 == 0, "whole waves");
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < NT / 64; ++w) r += red[w];
  __syncthreads();
  return r;
}

}  // namespace

// ---------------------------------------------------------------------------
// tape: the N slot pairs (s, x) of every site block <-> one tape column
// ---------------------------------------------------------------------------
__global__ void tape_copy_kernel(double* __restrict__ H, double* __restrict__ tape, long NP, int N, int s,
                                 int to_tape) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N * MB) return;
  const int x = idx / MB, lane = idx % MB, mb = blockIdx.y;
  double* h = H + hoff(mb, pair_off(s, x, N), lane, NP);
  double* tp = tape + ((((size_t)mb * N) + x) * MB + lane) * 2;
  if (to_tape)
    st2(tp, ld2(h));
  else
    st2(h, ld2(tp));
}

// packed tape of a cubic-symmetric state: one (F, r) pair per (slot x,
// orbit), read at the orbit's first site; layout [x][orbit][2]
__global__ void tape_pack_kernel(const double* __restrict__ H, double* __restrict__ tape,
                                 const int* __restrict__ orb_rep, int norb, long NP, int N, int s) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)norb * N) return;
  const int x = (int)(idx / norb), o = (int)(idx % norb);
  const int r = orb_rep[o];
  st2(tape + (size_t)idx * 2, ld2(H + hoff(r / MB, pair_off(s, x, N), r % MB, NP)));
}

// the primal column back from a packed tape: every site takes its orbit's pair
__global__ void tape_unpack_kernel(double* __restrict__ H, const double* __restrict__ tape,
                                   const int* __restrict__ site_orb, int V, int norb, long NP, int N, int s) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)V * N) return;
  const int x = (int)(idx / V), p = (int)(idx % V);
  st2(H + hoff(p / MB, pair_off(s, x, N), p % MB, NP), ld2(tape + ((size_t)x * norb + site_orb[p]) * 2));
}

// the three slot pairs of the initial rows (1,1), (2,1), (2,2) of every
// site, saved to or restored from buf[3][V][2]: with them the initial window
// [1,2] is rebuilt on a cleared history (host-snapshot checkpointing)
__global__ void adj_init_pairs_kernel(double* __restrict__ H, double* __restrict__ buf, long NP, int V, int N,
                                      int save) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= V) return;
  const int mb = p / MB, lane = p % MB;
  const int s1 = 0, s2 = 1 % N;
  const long pairs[3] = {pair_off(s1, s1, N), pair_off(s2, s1, N), pair_off(s2, s2, N)};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double* h = H + hoff(mb, pairs[k], lane, NP);
    double* b = buf + ((size_t)k * V + p) * 2;
    if (save)
      st2(b, ld2(h));
    else
      st2(h, ld2(b));
  }
}

// clear the adjoint of the slot pairs (s, x) and, with a tape column, put
// the primal column back
__global__ void adj_slide_kernel(double* __restrict__ H, double* __restrict__ Hbar, const double* __restrict__ tape,
                                 long NP, int N, int s) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N * MB) return;
  const int x = idx / MB, lane = idx % MB, mb = blockIdx.y;
  const size_t o = hoff(mb, pair_off(s, x, N), lane, NP);
  st2(Hbar + o, make_double2(0.0, 0.0));
  if (tape) st2(H + o, ld2(tape + ((((size_t)mb * N) + x) * MB + lane) * 2));
}

__global__ void adj_state_kernel(StepState* st, int t, int N) { *st = state_of_step(t, N); }

// dL/dF of F(T-1,T-1), F(T,T-1), F(T,T) for L = sum wgt n(T); n as
// createMeasurement (Measurement.jl:38-59) with its stability clamp
__global__ void adj_seed_kernel(const double* __restrict__ H, double* __restrict__ Hbar,
                                const double* __restrict__ wgt, long NP, int V, int N, int sT, int sTm1, double dt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= V) return;
  const int mb = p / MB, lane = p % MB;
  const double Ftt = H[hoff(mb, pair_off(sT, sT, N), lane, NP)];
  const double Fmt = H[hoff(mb, pair_off(sT, sTm1, N), lane, NP)];
  const double F = H[hoff(mb, pair_off(sTm1, sTm1, N), lane, NP)];
  const double ddF = (Ftt - 2.0 * Fmt + F) / (dt * dt);
  const double om2 = (ddF / F) > 0 ? ddF / F : 0.0;
  const bool stable = 0.25 * (dt * dt) * om2 < 1.0;
  double dn_dF = 0.0, dn_dd = 0.0;
  if (ddF * F > 0) {
    const double D = stable ? ddF * F - 0.25 * dt * dt * ddF * ddF : ddF * F;
    const double sq = sqrt(D);
    dn_dF = ddF / (2.0 * sq);
    dn_dd = stable ? (F - 0.5 * dt * dt * ddF) / (2.0 * sq) : F / (2.0 * sq);
  }
  const double w = wgt[p];
  const double g = w * dn_dd / (dt * dt);
  Hbar[hoff(mb, pair_off(sT, sT, N), lane, NP)] += g;
  Hbar[hoff(mb, pair_off(sT, sTm1, N), lane, NP)] += -2.0 * g;
  Hbar[hoff(mb, pair_off(sTm1, sTm1, N), lane, NP)] += w * dn_dF + g;
}

// ---------------------------------------------------------------------------
// Adjoint of the memory integrals (getRHS_F!, getRHS_r!, Evolution.jl:
// 766-782, 846-854) over the window [tmin, t], one workgroup per site block.
// Forward, per stored entry (a,b), a>b, with u = m w Sr, v = w SF, ms = m Sr:
//   R_F(a) -= u(b) F + v(b) rho,  R_F(b) -= u(a) F,  R_r(b) -= ms(a) rho,
// and (a,a): R_F(a) -= u(a) F.  With R_bar = dt^3 (adjoint of row t):
//   Sr_bar(z) = m(z) [ w(z) u_bar(z) + ms_bar(z) ],  SF_bar(z) = w(z) v_bar(z)
//   u_bar(z)  = -sum_b F(z,b) R_bar_F(b)           (b over [tmin, t])
//   v_bar(z)  = -sum_{a>z} rho(a,z) R_bar_F(a)
//   ms_bar(z) = -sum_{b<z} rho(z,b) R_bar_r(b)
// (adj_reduce_kernel, a read of the primal window only), and per entry
//   F_bar(a,b)   -= u(a) R_bar_F(b) + u(b) R_bar_F(a)
//   rho_bar(a,b) -= v(b) R_bar_F(a) + ms(a) R_bar_r(b).
// The entry updates are rank-2 per step.  They are deferred: the coefficient
// rings of up to AK consecutive reverse steps are kept and adj_apply_kernel
// adds all of them in one read-modify-write of the adjoint window; a row
// that is consumed in between (row t at reverse step t) gets the deferred
// terms added on the fly by adj_row_kernel.  Entries of a deferred step's
// update are those of its own window: times >= its tmin.
// Row t enters with R_bar_F(t) = dt^3 F_bar(t,t) (the diagonal's memory
// terms) and u = v = ms = 0; its own entries are cleared afterwards.
// The slot triangle is walked like the forward kernel: 8-lane workers take
// chunks of 4-column blocks (the forward's work items), two row batches in flight, column sums
// in registers, row sums in LDS.
// ---------------------------------------------------------------------------
constexpr int AM_THREADS = 256;
constexpr int AM_B = 1;  // reduce: rows per batch (two batches in flight)

struct SlotCoef {
  double u, v, ms, rf, rr, w, m;
};

__device__ __forceinline__ SlotCoef slot_coef(const AdjArgs& a, const double2* Ss, const double2* Rs, int x) {
  const StepState& st = a.st;
  const int N = a.N;
  const int rel = window_pos(x, st.s_min, N);
  const int tx = st.tmin + rel;
  SlotCoef k;
  k.w = rel == 0 ? 0.5 : 1.0;
  k.m = tx <= st.t - 2 ? 1.0 : 0.0;
  const double2 sg = tx <= st.t - 1 ? Ss[(size_t)x * MB] : make_double2(0.0, 0.0);
  const double2 rb = Rs[(size_t)x * MB];
  k.u = k.m * k.w * sg.y;
  k.v = k.w * sg.x;
  k.ms = k.m * sg.y;
  k.rf = rb.x;
  k.rr = rb.y;
  return k;
}

// coefficients of deferred step j at slot x whose time (in the current
// window) is tx; zero outside that step's window
struct PCoef {
  double u, v, ms, rf, rr;
};

__device__ __forceinline__ PCoef pend_coef(const AdjArgs& a, int j, int mb, int lane, int x, int tx) {
  // branch-free: the ring addresses are valid for every j < AK and slot x;
  // the values are selected (not multiplied) away, so stale ring contents
  // never reach the result
  const int k = a.pk[j];
  const size_t o = soff(mb, x, lane, a.N);
  const double2 sg = ld2(a.Sj[j] + o);
  const double2 rb = ld2(a.Rj[j] + o);
  const bool on = j < a.npend && tx >= a.ptmin[j];
  const double w = tx == a.ptmin[j] ? 0.5 : 1.0;
  const bool m = on && tx <= k - 2;
  const bool s = on && tx <= k - 1;
  PCoef c;
  c.u = m ? w * sg.y : 0.0;
  c.v = s ? w * sg.x : 0.0;
  c.ms = m ? sg.y : 0.0;
  c.rf = on ? rb.x : 0.0;
  c.rr = on ? rb.y : 0.0;
  return c;
}

// ---------------------------------------------------------------------------
// Adjoint of the diagonal and of the leapfrog row t (evolve_F!, evolve_r!,
// Evolution.jl:890-915, 971-974), one workgroup per site block.
//   F(t,t)  = (2-Om) F(t,t-1) - F(t,t-2) + dt^3 R_F(t)
//   F(t,b)  = (2-Om) F(t-1,b) - F(t-2,b) + dt^3 R_F(b),  b <= t-1
//   rho(t,b)= (2-Om) rho(t-1,b) - rho(t-2,b) + dt^3 R_r(b), b <= t-2
// with R_F(t) = -sum_{z<=t-2} w Sr(z) F(z,t) + sum_{z<=t-1} w SF(z) rho(z,t).
// The terms of R_F(t) that multiply row-t entries are folded in here; its
// self-energy terms are left to the memory-integral kernels through
// R_bar_F(t).
// ---------------------------------------------------------------------------
// threads per workgroup: 1024, or 512 where the coefficients of four or more
// deferred steps need more than 128 registers per lane
template <int NP>
constexpr int row_threads() {
  return NP > 3 ? 512 : 1024;
}

// NP: number of deferred steps after t (a compile-time count, so only
// their rings are read)
template <int NP>
__global__ __launch_bounds__(row_threads<NP>()) void adj_row_kernel(AdjArgs a) {
  constexpr int ROW_THREADS = row_threads<NP>();
  __shared__ double red[ROW_THREADS / 64];
  __shared__ double fl[2][MB];  // f_bar at t-1 and t-2
  const StepState& st = a.st;
  const int N = a.N;
  const int t = st.t;
  const int mb = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & (MB - 1);
  const int worker = tid >> 3;
  const int nworkers = ROW_THREADS >> 3;
  const int p = mb * MB + lane;
  const bool real = p < a.V;
  const double dt = a.dt, dt2 = dt * dt, dt3 = dt2 * dt;

  // Hartree adjoint of step t+1: M2_bar = dt^2 sum_p Om_bar(p)
  double Mb1 = 0.0;
  if (!a.first) {
    const double* Pn = a.P + (size_t)((t + 1) & 1) * a.NB;
    double acc = 0.0;
    for (int i = tid; i < a.NB; i += ROW_THREADS) acc += Pn[i];
    Mb1 = dt2 * block_sum<ROW_THREADS>(acc, red);
    if (mb == 0 && tid == 0) a.LP[0] += a.h_ivol * a.m2r[((t + 1) & 1) * 2 + 1] * Mb1;
  }
  const double M2 = a.m2r[(t & 1) * 2];
  const double Om = (a.kL2[p] + M2) * dt2;
  const double* Hs = a.H + hoff(mb, 0, lane, a.NP);
  double* Hbs = a.Hbar + hoff(mb, 0, lane, a.NP);
  const double2* Ss = reinterpret_cast<const double2*>(a.S + soff(mb, 0, lane, N));
  double2* Rs = reinterpret_cast<double2*>(a.Rb + soff(mb, 0, lane, N));
  const size_t pstride = (size_t)MB * 2;  // doubles per slot pair of one site block

  // deferred memory-integral terms of the steps after t that touch row t
  PCoef pt[NP > 0 ? NP : 1];
  double Fbtt = Hbs[pair_off(st.s_t, st.s_t, N) * pstride] + a.hl_ivol * Mb1;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    pt[j] = pend_coef(a, j, mb, lane, st.s_t, t);
    Fbtt -= pt[j].u * pt[j].rf;
  }
  const double RFt = dt3 * Fbtt;
  double omb = 0.0;
  for (int q = worker; q < st.L; q += nworkers) {
    const int tx = st.tmin + q;
    const int x = (tx - 1) % N;
    const double w = q == 0 ? 0.5 : 1.0;
    const double m = tx <= t - 2 ? 1.0 : 0.0;
    // every load of the iteration first: none of them is written by another
    // thread of this kernel
    double* e1p = Hbs + pair_off(st.s_tm1, x, N) * pstride;
    double* e2p = Hbs + pair_off(st.s_tm2, x, N) * pstride;
    const bool push1 = real && tx != t - 2, push2 = real && tx != t - 1;
    double2 hb = ld2(Hbs + pair_off(st.s_t, x, N) * pstride);
    const double2 h1 = ld2(Hs + pair_off(st.s_tm1, x, N) * pstride);
    const double2 sg = a.integral ? Ss[(size_t)x * MB] : make_double2(0.0, 0.0);
    double2 v1 = push1 ? ld2(e1p) : make_double2(0.0, 0.0);
    double2 v2 = push2 ? ld2(e2p) : make_double2(0.0, 0.0);
    PCoef px[NP > 0 ? NP : 1];
#pragma unroll
    for (int j = 0; j < NP; ++j) px[j] = pend_coef(a, j, mb, lane, x, tx);
    {
      // complete F_bar(t,tx), rho_bar(t,tx) (stored sign) with the deferred terms
      const double sgt = st.s_t >= x ? 1.0 : -1.0;
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        hb.x -= pt[j].u * px[j].rf + px[j].u * pt[j].rf;
        hb.y -= sgt * (px[j].v * pt[j].rf + pt[j].ms * px[j].rr);
      }
    }
    double fb = hb.x;
    if (tx == t - 1) fb += (2.0 - Om) * Fbtt;
    if (tx == t - 2) fb -= Fbtt;
    fb -= m * w * sg.y * RFt;
    double rp = 0.0;
    if (tx <= t - 2) rp = hb.y * (st.s_t >= x ? 1.0 : -1.0) - w * sg.x * RFt;
    const double s1 = st.s_tm1 >= x ? 1.0 : -1.0;
    omb -= h1.x * fb + h1.y * s1 * rp;
    Rs[(size_t)x * MB] = real ? make_double2(dt3 * fb, dt3 * rp) : make_double2(0.0, 0.0);
    // leapfrog pushes; the pair {t-1, t-2} is reached from b = t-1 and
    // b = t-2 and is updated once below
    if (push1) {
      v1.x += (2.0 - Om) * fb;
      if (tx <= t - 3) v1.y += s1 * (2.0 - Om) * rp;
      st2(e1p, v1);
    }
    if (push2) {
      v2.x -= fb;
      if (tx <= t - 3) v2.y -= (st.s_tm2 >= x ? 1.0 : -1.0) * rp;
      st2(e2p, v2);
    }
    if (tx == t - 1) fl[0][lane] = fb;
    if (tx == t - 2) fl[1][lane] = fb;
  }
  __syncthreads();
  if (tid < MB) {
    // F(t,t) depends on Om through (2-Om) F(t,t-1)
    omb -= Hs[pair_off(st.s_t, st.s_tm1, N) * pstride] * Fbtt;
    if (real) {
      double* e = Hbs + pair_off(st.s_tm1, st.s_tm2, N) * pstride;
      e[0] += (2.0 - Om) * fl[1][lane] - fl[0][lane];
      Rs[(size_t)st.s_t * MB] = make_double2(RFt, 0.0);
    } else {
      Rs[(size_t)st.s_t * MB] = make_double2(0.0, 0.0);
    }
  }
  const double tot = block_sum<ROW_THREADS>(real ? omb : 0.0, red);
  if (tid == 0) a.P[(size_t)(t & 1) * a.NB + mb] = tot;
}

// ---- reduce: Sigma_bar from one read of the primal window -----------------
struct RedCols {
  double rf[CB], rr[CB];
  double t1[CB], t2[CB], t3[CB];  // sum F rf(r); sum_{neg} P rr(r); sum_{not neg} P rf(r)
  unsigned phi;
};

struct RedBatch {
  double2 e[AM_B][CB];
  double2 rb[AM_B];
};

__device__ __forceinline__ void red_load(RedBatch& b, const double* Hs, const double2* Rs, long cbase, int c0, int r,
                                         int rlast) {
#pragma unroll
  for (int h = 0; h < AM_B; ++h) {
    const int rr = min(r + h, rlast);
    const long pb = cbase + (long)(rr - c0) * CB;
#pragma unroll
    for (int jc = 0; jc < CB; ++jc) b.e[h][jc] = ld2_nt(Hs + (size_t)(pb + jc) * MB * 2);
    b.rb[h] = Rs[(size_t)rr * MB];
  }
}

// one off-diagonal entry; n: time order reversed (time(r) < time(c))
__device__ __forceinline__ void red_pair(double2 e, bool n, double rfr, double rrr, RedCols& cs, int jc, double& s1,
                                         double& s2, double& s3) {
  const double Fv = e.x, Pv = e.y;
  const double nf = n ? 1.0 : 0.0, pf = 1.0 - nf;
  s1 += Fv * cs.rf[jc];
  cs.t1[jc] += Fv * rfr;
  s2 += pf * Pv * cs.rr[jc];
  cs.t3[jc] += pf * Pv * rfr;
  s3 += nf * Pv * cs.rf[jc];
  cs.t2[jc] += nf * Pv * rrr;
}

__device__ __forceinline__ void slot_wm(const StepState& st, int N, int x, double& w, double& m) {
  const int rel = window_pos(x, st.s_min, N);
  w = rel == 0 ? 0.5 : 1.0;
  m = st.tmin + rel <= st.t - 2 ? 1.0 : 0.0;
}

__device__ __forceinline__ void row_flush(const StepState& st, int N, int rr, double s1, double s2, double s3,
                                          double* accF, double* accR, int lane) {
  double w, m;
  slot_wm(st, N, rr, w, m);
  atomicAdd(&accR[rr * MB + lane], m * (-w * s1 - s2));
  atomicAdd(&accF[rr * MB + lane], w * s3);
}

__device__ __forceinline__ void red_batch(const AdjArgs& a, const RedBatch& b, int r, int r1, RedCols& cs,
                                          double* accF, double* accR, int lane) {
#pragma unroll
  for (int h = 0; h < AM_B; ++h) {
    const int rr = r + h;
    if (rr < r1) {
      const unsigned neg = rr >= a.st.s_min ? cs.phi : 0u;
      double s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
      for (int jc = 0; jc < CB; ++jc)
        red_pair(b.e[h][jc], (neg >> jc) & 1u, b.rb[h].x, b.rb[h].y, cs, jc, s1, s2, s3);
      row_flush(a.st, a.N, rr, s1, s2, s3, accF, accR, lane);
    }
  }
}

__global__ __launch_bounds__(AM_THREADS, 2) void adj_reduce_kernel(AdjArgs a) {
  extern __shared__ double smem[];
  const int N = a.N;
  double* accF = smem;           // SigF_bar per slot, lane
  double* accR = accF + N * MB;  // Sigr_bar
  __shared__ double red[AM_THREADS / 64];
  __shared__ int next_item;
  const StepState& st = a.st;
  const int t = st.t;
  const int Rv = t > N ? N : t;  // slots in use: the window [tmin, t]
  const int smin = st.s_min;
  const int mb = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & (MB - 1);
  const int worker = tid >> 3;
  const int nworkers = AM_THREADS >> 3;
  for (int i = tid; i < N * MB; i += AM_THREADS) {
    accF[i] = 0.0;
    accR[i] = 0.0;
  }
  if (tid == 0) next_item = nworkers;
  __syncthreads();
  const double* Hs = a.H + hoff(mb, 0, lane, a.NP);
  const double2* Ss = reinterpret_cast<const double2*>(a.S + soff(mb, 0, lane, N));
  const double2* Rs = reinterpret_cast<const double2*>(a.Rb + soff(mb, 0, lane, N));

  int item = worker;
  while (item < a.nitems) {
    const int2 jk = a.items[item];
    const int c0 = jk.x * CB;
    const int r0 = c0 + jk.y * MEMINT_ROWS;
    const int r1 = min(r0 + MEMINT_ROWS, Rv);
    if (c0 < Rv && r0 < r1) {
      const long cbase = colblock_offset(jk.x, N);
      RedCols cs;
      cs.phi = 0;
#pragma unroll
      for (int jc = 0; jc < CB; ++jc) {
        const int c = c0 + jc;
        const double2 rb = Rs[(size_t)min(c, Rv - 1) * MB];
        const bool ok = c < Rv;
        cs.rf[jc] = ok ? rb.x : 0.0;
        cs.rr[jc] = ok ? rb.y : 0.0;
        cs.t1[jc] = cs.t2[jc] = cs.t3[jc] = 0.0;
        cs.phi |= (c < smin ? 1u : 0u) << jc;
      }
      int r = r0;
      // rows crossing the block diagonal: columns c0 .. r
      for (; r < r1 && r < c0 + CB; ++r) {
        const double2 rbr = Rs[(size_t)r * MB];
        const unsigned neg = r >= smin ? cs.phi : 0u;
        const long pb = cbase + (long)(r - c0) * CB;
        double s1 = 0.0, s2 = 0.0, s3 = 0.0;
        for (int jc = 0; jc < CB; ++jc) {
          const int c = c0 + jc;
          if (c > r) break;
          const double2 e = ld2(Hs + (size_t)(pb + jc) * MB * 2);
          if (c == r)
            s1 += e.x * rbr.x;
          else
            red_pair(e, (neg >> jc) & 1u, rbr.x, rbr.y, cs, jc, s1, s2, s3);
        }
        row_flush(st, N, r, s1, s2, s3, accF, accR, lane);
      }
      if (r < r1) {
        const int rlast = r1 - 1;
        RedBatch A, B;
        red_load(A, Hs, Rs, cbase, c0, r, rlast);
        while (true) {
          red_load(B, Hs, Rs, cbase, c0, r + AM_B, rlast);
          red_batch(a, A, r, r1, cs, accF, accR, lane);
          r += AM_B;
          if (r >= r1) break;
          red_load(A, Hs, Rs, cbase, c0, r + AM_B, rlast);
          red_batch(a, B, r, r1, cs, accF, accR, lane);
          r += AM_B;
          if (r >= r1) break;
        }
      }
#pragma unroll
      for (int jc = 0; jc < CB; ++jc) {
        const int c = c0 + jc;
        if (c < Rv) {
          double w, m;
          slot_wm(st, N, c, w, m);
          atomicAdd(&accR[c * MB + lane], m * (-w * cs.t1[jc] + cs.t2[jc]));
          atomicAdd(&accF[c * MB + lane], -w * cs.t3[jc]);
        }
      }
    }
    int nxt = 0;
    if (lane == 0) nxt = atomicAdd(&next_item, 1);
    item = __shfl(nxt, tid & ~(MB - 1) & 63);
  }
  __syncthreads();

  // Sigma_bar of the live times t' in [tmin, t-1]; dL/dlambda through the
  // lambda^2 prefactors of the sunset
  const int p = mb * MB + lane;
  const bool real = p < a.V;
  double2* Sb = reinterpret_cast<double2*>(a.Sbar + soff(mb, 0, lane, N));
  double lam = 0.0;
  for (int q = worker; q < st.L; q += nworkers) {
    const int x = (st.tmin + q - 1) % N;
    const double2 sb = make_double2(accF[x * MB + lane], accR[x * MB + lane]);
    Sb[(size_t)x * MB] = real ? sb : make_double2(0.0, 0.0);
    const double2 sg = Ss[(size_t)x * MB];
    if (real) lam += sg.x * sb.x + sg.y * sb.y;
  }
  const double tot = block_sum<AM_THREADS>(lam, red);
  if (tid == 0) a.LP[mb] += a.two_over_lambda * tot;
}

// ---- apply: the deferred entry updates of npend steps ----------------------
// The triangle of every site block is cut into chunks of AP_ROWS slot rows,
// one workgroup per (site block, chunk).  The workgroup stages its rows'
// coefficients (u = w ms, rf, ms, v, rr per deferred step and lane) in LDS.
// Each wave then takes panels of 8 columns (two column blocks, one column per
// 8-lane worker) and walks the chunk's rows with all its workers on the same
// row, so the staged row coefficients are read as broadcasts.  A worker keeps
// its column's coefficients in registers, sums the AK steps' terms of an
// entry and adds them in one read-modify-write of the adjoint line, with
// two batches of AP_D rows in flight:
//   F_bar(a,b) -= sum_j u_j(a) rf_j(b) + rf_j(a) u_j(b)     (half of it at a = b)
//   rho_bar    -= sum_j v_j(b) rf_j(a) + ms_j(a) rr_j(b)   (a later than b)
// The stored pair (row, column) has the row later in time unless the column
// lies before s_min and the row does not (the ring has wrapped); then the
// roles swap and the stored rho changes sign.
// Workgroups are dealt to the 8 XCDs in turn; the chunks of one site block
// are consecutive workgroups of one XCD, so the coefficient rings that every
// chunk reads for its columns come from HBM about once and then from that
// XCD's L2.
constexpr int AP_THREADS = 256;
constexpr int AP_ROWS = 28;  // rows per chunk: 70 KB of staged coefficients
constexpr int AP_D = 4;      // rows per batch
constexpr int AP_Q = 5;      // staged values per (row, step, lane)

// staged coefficients of (row, step j): (u, rf) of the 8 lanes, (ms, v) of
// the 8 lanes, rr of the 8 lanes
__device__ __forceinline__ int ap_base(int row, int j) { return (row * AK + j) * AP_Q * MB; }

__host__ __device__ inline int apply_chunks(int Rv) { return (Rv + AP_ROWS - 1) / AP_ROWS; }

struct ApCol {
  double cf[AK], uc[AK], cv[AK], crr[AK], cms[AK];
};

// The adjoint lines are read and written through a buffer resource over the
// site block's region: a lane whose entry is not updated (above the
// diagonal, or a row past the chunk in the last batch) gets an offset beyond
// the resource's size, and the hardware then returns zeros for its load and
// drops its store, so masked entries cost no memory traffic and the loop
// needs no branch.  Dword 3 of the resource: 32-bit data format, the value
// for the gfx9 family.
constexpr unsigned AP_OFF_NONE = 0x80000000u;
constexpr int AP_NT = 2;  // cache policy of the adjoint lines: non-temporal (streamed once per call)
typedef unsigned kb_u4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t ap_rsrc(double* base, long NP) {
  return __builtin_amdgcn_make_buffer_rsrc(base, 0, (int)(NP * MB * 16), 0x00020000);
}

// byte offset of this worker's line in row rr, or AP_OFF_NONE
__device__ __forceinline__ unsigned ap_off(unsigned col_off, int c, int c0, int rr, int rlast) {
  return (rr <= rlast && c <= rr) ? col_off + (unsigned)(rr - c0) * (CB * MB * 16) : AP_OFF_NONE;
}

__device__ __forceinline__ void ap_load(double2 (&e)[AP_D], __amdgpu_buffer_rsrc_t rs, unsigned col_off, int c, int c0,
                                        int r, int rlast) {
#pragma unroll
  for (int d = 0; d < AP_D; ++d) {
    const kb_u4v v = __builtin_amdgcn_raw_buffer_load_b128(rs, ap_off(col_off, c, c0, r + d, rlast), 0, AP_NT);
    e[d] = __builtin_bit_cast(double2, v);
  }
}

__device__ __forceinline__ void ap_rows(double2 (&e)[AP_D], __amdgpu_buffer_rsrc_t rs, unsigned col_off,
                                        const double* rco, const ApCol& k, int c, int c0, int r, int rlast, int R0,
                                        bool cwrap, int smin, int lane) {
#pragma unroll
  for (int d = 0; d < AP_D; ++d) {
    const int rc = min(r + d, rlast);
    const double* rb = rco + ap_base(rc - R0, 0);
    double ax = 0.0, yp = 0.0, yn = 0.0;
#pragma unroll
    for (int j = 0; j < AK; ++j) {
      const double2 A = *reinterpret_cast<const double2*>(rb + j * AP_Q * MB + 2 * lane);           // u, rf
      const double2 B = *reinterpret_cast<const double2*>(rb + j * AP_Q * MB + 2 * MB + 2 * lane);  // ms, v
      const double R = rb[j * AP_Q * MB + 4 * MB + lane];                                           // rr
      ax = fma(A.x, k.cf[j], ax);
      ax = fma(k.uc[j], A.y, ax);
      yp = fma(k.cv[j], A.y, yp);
      yp = fma(B.x, k.crr[j], yp);
      yn = fma(B.y, k.cf[j], yn);
      yn = fma(k.cms[j], R, yn);
    }
    // c == rc: half of the F term, no rho term; c > rc: the store is dropped
    const bool nf = cwrap && rc >= smin;
    e[d].x -= (c < rc ? 1.0 : 0.5) * ax;
    e[d].y += c < rc ? (nf ? yn : -yp) : 0.0;
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(kb_u4v, e[d]), rs,
                                           ap_off(col_off, c, c0, r + d, rlast), 0, AP_NT);
  }
}

__global__ __launch_bounds__(AP_THREADS, 2) void adj_apply_kernel(AdjArgs a) {
  // AP_ROWS x AK x AP_Q x MB staged values, then the panel counter; no
  // static LDS, so the staged pairs start 16-byte aligned
  extern __shared__ double rco[];
  int& next_item = *reinterpret_cast<int*>(rco + AP_ROWS * AK * AP_Q * MB);
  const StepState& st = a.st;
  const int N = a.N;
  const int Rv = st.t > N ? N : st.t;
  const int smin = st.s_min;
  const int nchunk = apply_chunks(Rv);
  const int g = blockIdx.x;
  const int q = g >> 3;
  const int mbg = q / nchunk;
  const int mb = mbg * 8 + (g & 7);
  if (mb >= a.NB) return;
  const int chunk = nchunk - 1 - (q - mbg * nchunk);  // largest chunk first
  const int R0 = chunk * AP_ROWS;
  const int R1 = min(R0 + AP_ROWS, Rv);
  const int tid = threadIdx.x;
  const int lane = tid & (MB - 1);
  for (int i = tid >> 3; i < (R1 - R0) * AK; i += AP_THREADS >> 3) {
    const int row = i / AK, j = i - row * AK;
    const int r = R0 + row;
    const PCoef k = pend_coef(a, j, mb, lane, r, st.tmin + window_pos(r, smin, N));
    double* b = rco + ap_base(row, j);
    b[2 * lane] = k.u;
    b[2 * lane + 1] = k.rf;
    b[2 * MB + 2 * lane] = k.ms;
    b[2 * MB + 2 * lane + 1] = k.v;
    b[4 * MB + lane] = k.rr;
  }
  if (tid == 0) next_item = AP_THREADS / 64;
  __syncthreads();
  const int w = (tid & 63) >> 3;  // this worker's column in the panel
  const __amdgpu_buffer_rsrc_t rs = ap_rsrc(a.Hbar + hoff(mb, 0, 0, a.NP), a.NP);
  const int npan = (R1 + 7) / 8;
  const int rlast = R1 - 1;
  int p = tid >> 6;
  while (p < npan) {
    const int c = 8 * p + w;
    const int J = c / CB;
    const int c0 = J * CB;
    if (c0 < R1) {  // the column block has rows in this chunk
      const int cc = min(c, Rv - 1);
      const int txc = st.tmin + window_pos(cc, smin, N);
      const bool ok = c < Rv;
      ApCol k;
#pragma unroll
      for (int j = 0; j < AK; ++j) {
        const PCoef q = pend_coef(a, j, mb, lane, cc, txc);
        k.cf[j] = ok ? q.rf : 0.0;
        k.uc[j] = ok ? q.u : 0.0;
        k.cv[j] = ok ? q.v : 0.0;
        k.crr[j] = ok ? q.rr : 0.0;
        k.cms[j] = ok ? q.ms : 0.0;
      }
      const unsigned col_off = (unsigned)((colblock_offset(J, N) + (c - c0)) * MB + lane) * 16u;
      double2 X[AP_D], Y[AP_D];
      int r = max(R0, 8 * p);
      ap_load(X, rs, col_off, c, c0, r, rlast);
      // the scheduling barriers keep each batch's loads ahead of the other
      // batch's arithmetic
      while (true) {
        ap_load(Y, rs, col_off, c, c0, r + AP_D, rlast);
        __builtin_amdgcn_sched_barrier(0);
        ap_rows(X, rs, col_off, rco, k, c, c0, r, rlast, R0, c < smin, smin, lane);
        r += AP_D;
        if (r >= R1) break;
        ap_load(X, rs, col_off, c, c0, r + AP_D, rlast);
        __builtin_amdgcn_sched_barrier(0);
        ap_rows(Y, rs, col_off, rco, k, c, c0, r, rlast, R0, c < smin, smin, lane);
        r += AP_D;
        if (r >= R1) break;
      }
    }
    int nx = 0;
    if ((tid & 63) == 0) nx = atomicAdd(&next_item, 1);
    p = __shfl(nx, 0);
  }
}

size_t apply_lds() { return sizeof(double) * (size_t)AP_ROWS * AK * AP_Q * MB + 16; }

// the Hartree adjoint of step 3 and the initial-row gradients:
// F(1,1) = F(2,1) = (n+1/2)/w, F(2,2) = (n+1/2)(w dt^2 + 1/w)
__global__ __launch_bounds__(1024) void adj_pending_kernel(const double* __restrict__ P, int NB, double dt,
                                                           double h_ivol, const double* __restrict__ m2r,
                                                           double* __restrict__ LP, double* __restrict__ Mb_out) {
  __shared__ double red[1024 / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < NB; i += blockDim.x) acc += P[(size_t)(3 & 1) * NB + i];
  const double Mb = dt * dt * block_sum<1024>(acc, red);
  if (threadIdx.x == 0) {
    Mb_out[0] = Mb;
    LP[0] += h_ivol * m2r[(3 & 1) * 2 + 1] * Mb;
  }
}

__global__ void adj_final_kernel(const double* __restrict__ Hbar, long NP, int V, int N, const double* kL2,
                                 double initmass, int have_mass, double dt, double hl_ivol,
                                 const double* __restrict__ Mb, double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= V) return;
  const int mb = p / MB, lane = p % MB;
  const int s1 = 0, s2 = 1 % N;
  const double f11 = Hbar[hoff(mb, pair_off(s1, s1, N), lane, NP)];
  const double f21 = Hbar[hoff(mb, pair_off(s2, s1, N), lane, NP)];
  const double f22 = Hbar[hoff(mb, pair_off(s2, s2, N), lane, NP)] + (Mb ? hl_ivol * Mb[0] : 0.0);
  out[p] = f11;
  out[V + p] = f21;
  out[2 * V + p] = f22;
  if (have_mass) {
    const double w = sqrt(kL2[p] + initmass * initmass);
    out[3 * V + p] = f11 / w + f21 / w + f22 * (w * dt * dt + 1.0 / w);
  }
}

}  // namespace qftkb
#pragma GCC visibility pop

// ===========================================================================
// host side
// ===========================================================================
#pragma GCC visibility push(hidden)
namespace qftkb_host {

struct AdjState {
  int C = 0;          // checkpoint spacing in steps; 0: tape every dropped column
  int max_steps = 0;  // largest T that may be recorded
  double* Hbar = nullptr;
  double* Sring[AK] = {};  // Sigma(k-1, .) of reverse step k, at k % AK
  double* Rring[AK] = {};  // R_bar of reverse step k, at k % AK
  double* Sbar = nullptr;
  double2* X = nullptr;
  double* P = nullptr;    // [2][NB]
  double* LP = nullptr;   // [NB]
  double* m2r = nullptr;  // [2][2]
  double* Mb = nullptr;   // [1]
  double* out = nullptr;  // 4 x V
  double* tape = nullptr;
  int tape_cols = 0;
  size_t col_doubles = 0;  // one tape column
  // packed tape (cubic-symmetric states): one value per orbit and slot
  bool packed = false;
  int norb = 0;
  int* orb_rep = nullptr;   // orbit -> first site
  int* site_orb = nullptr;  // site -> orbit
  std::vector<double*> ckpt;
  // host-snapshot checkpointing (hwin >= 0): up to hwin windows at segment
  // boundaries in host memory (hsnap[i] holds boundary hseg[i], -1: free),
  // copied through two pinned staging buffers (pinned memory is scarce: a
  // window is larger than what the driver pins), the initial window as its
  // three rows on the device (initbuf); plan: the boundaries the recording
  // run keeps
  int hwin = -1;
  std::vector<char*> hsnap;
  char* stage[2] = {};
  hipEvent_t stage_ev[2] = {};
  std::vector<int> hseg;
  std::vector<int> plan;
  std::vector<int> rvF, rvG;  // best first snapshot (segments) per (n, free slots)
  int rv_w = 1;
  double* initbuf = nullptr;
  size_t host_bytes = 0;
  bool progress = false;  // QFTKB_PROGRESS set: one stderr line per recomputed segment
  size_t hbytes = 0;
  size_t bytes = 0;
  long long recomputed = 0;
  int stop_after = 0;  // tests: end the next sweep after this reverse step
};

namespace {

int amalloc(kb_ctx* c, AdjState* s, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 8);
  if (e != hipSuccess) {
    *p = nullptr;
    return seterr(c, KB_ERR_OOM, "adjoint: hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  }
  s->bytes += bytes;
  return KB_OK;
}

void free_state(AdjState* s) {
  if (!s) return;
  for (void* p : {(void*)s->Hbar, (void*)s->Sbar, (void*)s->X, (void*)s->P, (void*)s->LP, (void*)s->m2r,
                  (void*)s->Mb, (void*)s->out, (void*)s->tape, (void*)s->orb_rep, (void*)s->site_orb,
                  (void*)s->initbuf})
    if (p) (void)hipFree(p);
  for (char* p : s->hsnap) std::free(p);
  for (int k = 0; k < 2; ++k) {
    if (s->stage[k]) (void)hipHostFree(s->stage[k]);
    if (s->stage_ev[k]) (void)hipEventDestroy(s->stage_ev[k]);
  }
  for (int j = 0; j < AK; ++j) {
    if (s->Sring[j]) (void)hipFree(s->Sring[j]);
    if (s->Rring[j]) (void)hipFree(s->Rring[j]);
  }
  for (double* p : s->ckpt)
    if (p) (void)hipFree(p);
  delete s;
}

// first step of the segment that contains step t (checkpoint at seg - 1)
inline int seg_first(int t, int C) { return 3 + C * ((t - 3) / C); }

inline int tape_index(const kb_ctx* c, int t) {
  const AdjState* s = c->adj;
  return s->C > 0 ? t - seg_first(t, s->C) : t - c->N - 1;
}

int save_column(kb_ctx* c, int t) {
  AdjState* s = c->adj;
  const StepState st = state_of_step(t, c->N);
  if (st.dead < 0) return KB_OK;
  const int k = tape_index(c, t);
  if (k < 0 || k >= s->tape_cols) return seterr(c, KB_ERR_STATE, "adjoint: tape column %d out of range", k);
  if (s->packed) {
    const long n = (long)s->norb * c->N;
    hipLaunchKernelGGL(tape_pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->H,
                       s->tape + (size_t)k * s->col_doubles, s->orb_rep, s->norb, c->NP, c->N, st.s_t);
    LAUNCHCHK(c);
    return KB_OK;
  }
  const dim3 grid((c->N * MB + 255) / 256, c->NB);
  hipLaunchKernelGGL(tape_copy_kernel, grid, dim3(256), 0, c->stream, c->H, s->tape + (size_t)k * s->col_doubles,
                     c->NP, c->N, st.s_t, 1);
  LAUNCHCHK(c);
  return KB_OK;
}

AdjArgs adj_args(kb_ctx* c, int t, int T) {
  AdjState* s = c->adj;
  AdjArgs a{};
  a.H = c->H;
  a.Hbar = s->Hbar;
  a.S = s->Sring[t % AK];
  a.Rb = s->Rring[t % AK];
  a.Sbar = s->Sbar;
  a.kL2 = c->kL2;
  a.m2r = s->m2r;
  a.P = s->P;
  a.LP = s->LP;
  a.items = c->items;
  a.nitems = c->nitems;
  a.NP = c->NP;
  a.N = c->N;
  a.V = c->V;
  a.NB = c->NB;
  a.dt = c->p.dt;
  const double h = (double)(c->p.ONgroup + 2) / (6.0 * (double)c->p.ONgroup);
  a.h_ivol = h / (double)c->V;
  a.hl_ivol = c->p.Lambda * h / (double)c->V;
  a.two_over_lambda = c->p.Lambda != 0.0 ? 2.0 / c->p.Lambda : 0.0;
  a.integral = sunset_expansion(c->p.pexp);
  a.first = t == T;
  a.st = state_of_step(t, c->N);
  return a;
}

int launch_sigma_adjoint(kb_ctx* c) {
  AdjState* s = c->adj;
  const double ivol = 1.0 / (double)c->V;
  const double lam = c->p.Lambda, ON = (double)c->p.ONgroup;
  SigmaArgs a{};
  a.H = c->H;
  a.W = c->W;
  a.tw = c->tw;
  a.plan = c->plan;
  a.NP = c->NP;
  a.N = c->N;
  a.V = c->V;
  a.n = c->p.Nx;
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.st = c->st_dev;
  a.preF = ivol * ivol * ivol * (-(lam * lam) * (ON + 2.0) / (18.0 * ON * ON));
  a.preR = ivol * ivol * ivol * (-(lam * lam) * (ON + 2.0) / (6.0 * ON * ON));
  a.X = s->X;
  a.Sbar = s->Sbar;
  a.Hbar = s->Hbar;
  a.classical = c->p.pexp == KB_PEXP_LAMBDA_NLO_CLASSICAL;
  a.P = c->fftP;
  a.npencils = c->V / c->p.Nx;
  hipError_t e = sigma_planes_supported(c->p.Nx, c->p.sdim) ? launch_sigma_planes_adjoint(a, c->N - 1, c->stream)
                                                            : launch_sunset_adjoint_pencils(a, c->N - 1, c->stream);
  if (e != hipSuccess) return seterr(c, KB_ERR_HIP, "adjoint self-energy launch failed: %s", hipGetErrorString(e));
  return KB_OK;
}

// adj_row_kernel with its compile-time count of deferred steps (0 .. AK-1)
template <int NP>
void launch_row(int np, dim3 grid, hipStream_t st, const AdjArgs& a) {
  if constexpr (NP + 1 < AK) {
    if (np > NP) return launch_row<NP + 1>(np, grid, st, a);
  }
  hipLaunchKernelGGL(adj_row_kernel<NP>, grid, dim3(row_threads<NP>()), 0, st, a);
}

// deferred steps k = k0, k0+1, .. k0+n-1 into the argument block
void set_pending(kb_ctx* c, AdjArgs& a, int k0, int n) {
  AdjState* s = c->adj;
  a.npend = n;
  for (int j = 0; j < AK; ++j) {
    // slots past n point at valid rings too; the kernels mask them out
    const int k = k0 + j;
    a.pk[j] = k;
    a.ptmin[j] = k > c->N ? k - c->N + 1 : 1;
    a.Sj[j] = s->Sring[k % AK];
    a.Rj[j] = s->Rring[k % AK];
  }
}

// reverse step t: the primal window of step t is in H.  npend: steps
// t+1 .. t+npend whose entry updates are deferred; apply: add them and this
// step's to the adjoint window now.
int reverse_step(kb_ctx* c, int t, int T, int npend, bool apply) {
  AdjState* s = c->adj;
  const int N = c->N;
  const bool nlo = sunset_expansion(c->p.pexp);
  AdjArgs a = adj_args(c, t, T);
  int rc;
  {
    StageTimer tm(c, 4);
    hipLaunchKernelGGL(adj_state_kernel, dim3(1), dim3(1), 0, c->stream, c->st_dev, t, N);
    LAUNCHCHK(c);
    if (nlo && (rc = launch_sigma(c, s->X, s->Sring[t % AK]))) return rc;
    const double H = c->p.Lambda * (double)(c->p.ONgroup + 2) / (6.0 * (double)c->p.ONgroup);
    // M^2(t) and sum_p F(t-1,t-1;p) of the step being reversed (st_dev holds its state)
    double* m2t = s->m2r + (t & 1) * 2;
    hipLaunchKernelGGL(hartree_step_kernel, dim3(hartree_groups(c->V)), dim3(HARTREE_THREADS), 0, c->stream, c->H,
                       c->NP, c->V, N, c->st_dev, H / (double)c->Vl, c->p.Mass2, nullptr, m2t, c->hpart, m2t + 1);
    LAUNCHCHK(c);
    set_pending(c, a, t + 1, nlo ? npend : 0);
    launch_row<0>(a.npend, dim3(c->NB), c->stream, a);
    LAUNCHCHK(c);
  }
  if (nlo) {
    StageTimer tm(c, 3);
    hipLaunchKernelGGL(adj_reduce_kernel, dim3(c->NB), dim3(AM_THREADS), memint_lds(N), c->stream, a);
    LAUNCHCHK(c);
  }
  {
    StageTimer tm(c, 4);
    if (nlo && (rc = launch_sigma_adjoint(c))) return rc;
  }
  if (nlo && apply) {
    StageTimer tm(c, 3);
    set_pending(c, a, t, npend + 1);
    const int Rv = t > N ? N : t;
    hipLaunchKernelGGL(adj_apply_kernel, dim3((unsigned)((c->NB + 7) / 8 * 8 * apply_chunks(Rv))), dim3(AP_THREADS),
                       apply_lds(), c->stream, a);
    LAUNCHCHK(c);
  }
  {
    StageTimer tm(c, 4);
    const double* col = nullptr;
    if (a.st.dead >= 0) col = s->tape + (size_t)tape_index(c, t) * s->col_doubles;
    const dim3 grid((N * MB + 255) / 256, c->NB);
    hipLaunchKernelGGL(adj_slide_kernel, grid, dim3(256), 0, c->stream, c->H, s->Hbar, s->packed ? nullptr : col,
                       c->NP, N, a.st.s_t);
    LAUNCHCHK(c);
    if (s->packed && col) {
      const long n = (long)c->V * N;
      hipLaunchKernelGGL(tape_unpack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->H, col,
                         s->site_orb, c->V, s->norb, c->NP, N, a.st.s_t);
      LAUNCHCHK(c);
    }
  }
  return KB_OK;
}

// ---- host-snapshot checkpointing --------------------------------------------
// Segments of C steps; boundary b is the window after step 2 + b C.  The
// reverse sweep takes the segments from the last to the first and needs the
// window at each one's start.  F(n, f): the least segments advanced to
// reverse n segments whose start window is held, with f free snapshot slots:
//   F(1, f) = 0,  F(n, 0) = n (n - 1) / 2,
//   F(n, f) = min over m of  m + F(n - m, f - 1) + F(m, f)
// (advance m segments and keep that window, reverse the n - m segments
// above it with one slot less, free it, reverse the m below).  This is the
// binomial checkpointing problem (Griewank's revolve), whose optimum is
// r n - binomial(f + 2 + r - 1, f + 2) for binomial(f + r, f + 1) < n <=
// binomial(f + 1 + r, f + 1).  G(n, f) is the same for the recording run,
// which advances over every boundary anyway and tapes the last segment:
//   G(1, f) = 0,  G(n, f) = min(F(n - 1, f), min over m of G(n - m, f - 1) + F(m, f)).
// The tables keep the best m (0: keep no window).
void revolve_tables(AdjState* s, int nseg, int slots) {
  const size_t w = (size_t)slots + 1;
  s->rv_w = (int)w;
  s->rvF.assign((size_t)(nseg + 1) * w, 0);
  s->rvG.assign((size_t)(nseg + 1) * w, 0);
  std::vector<long long> F((size_t)(nseg + 1) * w, 0), G((size_t)(nseg + 1) * w, 0);
  for (int n = 2; n <= nseg; ++n)
    for (int f = 0; f <= slots; ++f) {
      long long best = (long long)n * (n - 1) / 2;
      int bm = 0;
      for (int m = 1; f > 0 && m < n; ++m) {
        const long long c = m + F[(size_t)(n - m) * w + f - 1] + F[(size_t)m * w + f];
        if (c < best) best = c, bm = m;
      }
      F[(size_t)n * w + f] = best;
      s->rvF[(size_t)n * w + f] = bm;
      long long gbest = F[(size_t)(n - 1) * w + f];
      int gm = 0;
      for (int m = 1; f > 0 && m < n; ++m) {
        const long long c = G[(size_t)(n - m) * w + f - 1] + F[(size_t)m * w + f];
        if (c < gbest) gbest = c, gm = m;
      }
      G[(size_t)n * w + f] = gbest;
      s->rvG[(size_t)n * w + f] = gm;
    }
}

// boundaries the recording run keeps: the right spine of G over all segments
std::vector<int> host_plan(const AdjState* s, int nseg, int slots) {
  std::vector<int> plan;
  int cur = 0, n = nseg;
  for (int f = slots; f > 0 && n >= 2; --f) {
    const int m = s->rvG[(size_t)n * s->rv_w + f];
    if (m == 0) break;
    cur += m;
    n -= m;
    plan.push_back(cur);
  }
  return plan;
}

int free_slot(const AdjState* s) {
  for (size_t i = 0; i < s->hseg.size(); ++i)
    if (s->hseg[i] < 0) return (int)i;
  return -1;
}

// staged copies between the device window and a host snapshot: chunks of
// STAGE_BYTES through two pinned buffers, the host side copied by several
// threads while the next chunk crosses PCIe
constexpr size_t STAGE_BYTES = size_t(1) << 30;
constexpr int COPY_THREADS = 8;

void parallel_copy(char* dst, const char* src, size_t n) {
  std::vector<std::thread> th;
  const size_t part = (n + COPY_THREADS - 1) / COPY_THREADS;
  for (int k = 0; k < COPY_THREADS; ++k) {
    const size_t o = (size_t)k * part;
    if (o >= n) break;
    const size_t len = std::min(part, n - o);
    th.emplace_back([=] { std::memcpy(dst + o, src + o, len); });
  }
  for (auto& t : th) t.join();
}

int snapshot_out(kb_ctx* c, char* snap) {  // device window -> host
  AdjState* s = c->adj;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const char* H = reinterpret_cast<const char*>(c->H);
  const size_t nch = (s->hbytes + STAGE_BYTES - 1) / STAGE_BYTES;
  auto len = [&](size_t i) { return std::min(STAGE_BYTES, s->hbytes - i * STAGE_BYTES); };
  for (size_t i = 0; i <= nch; ++i) {
    if (i < nch) {
      HIPCHK(c, hipMemcpyAsync(s->stage[i % 2], H + i * STAGE_BYTES, len(i), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipEventRecord(s->stage_ev[i % 2], c->stream));
    }
    if (i > 0) {  // chunk i-1 landed: copy it out while chunk i crosses
      HIPCHK(c, hipEventSynchronize(s->stage_ev[(i - 1) % 2]));
      parallel_copy(snap + (i - 1) * STAGE_BYTES, s->stage[(i - 1) % 2], len(i - 1));
    }
  }
  return KB_OK;
}

int snapshot_in(kb_ctx* c, const char* snap) {  // host -> device window
  AdjState* s = c->adj;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  char* H = reinterpret_cast<char*>(c->H);
  const size_t nch = (s->hbytes + STAGE_BYTES - 1) / STAGE_BYTES;
  for (size_t i = 0; i < nch; ++i) {
    const size_t n = std::min(STAGE_BYTES, s->hbytes - i * STAGE_BYTES);
    if (i >= 2) HIPCHK(c, hipEventSynchronize(s->stage_ev[i % 2]));  // its last upload is done
    parallel_copy(s->stage[i % 2], snap + i * STAGE_BYTES, n);
    HIPCHK(c, hipMemcpyAsync(H + i * STAGE_BYTES, s->stage[i % 2], n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipEventRecord(s->stage_ev[i % 2], c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return KB_OK;
}

// copy the window (boundary b) into a free host slot
int store_snapshot(kb_ctx* c, int b) {
  AdjState* s = c->adj;
  const int i = free_slot(s);
  if (i < 0) return seterr(c, KB_ERR_STATE, "adjoint: no free host snapshot slot");
  int rc;
  if ((rc = snapshot_out(c, s->hsnap[i]))) return rc;
  s->hseg[i] = b;
  return KB_OK;
}

// run forward steps from .. to (no tape); the window of step from - 1 is in H
int advance_steps(kb_ctx* c, int from, int to) {
  int rc;
  for (int t = from; t <= to; ++t) {
    advance_state(c->st, c->N);
    if ((rc = enqueue_step(c))) return rc;
    c->adj->recomputed += 1;
  }
  return KB_OK;
}

// put the window of boundary k into H from the nearest snapshot below it,
// keeping new snapshots on the way by the binomial schedule
int host_window(kb_ctx* c, int k) {
  AdjState* s = c->adj;
  const int C = s->C;
  int j = 0, slot = -1;
  for (size_t i = 0; i < s->hseg.size(); ++i) {
    if (s->hseg[i] > k) s->hseg[i] = -1;  // its segments are reversed
    if (s->hseg[i] >= 0 && s->hseg[i] >= j) {
      j = s->hseg[i];
      slot = (int)i;
    }
  }
  if (slot >= 0) {
    int rc;
    if ((rc = snapshot_in(c, s->hsnap[slot]))) return rc;
  } else {
    // the initial window: the three initial rows on a cleared history
    HIPCHK(c, hipMemsetAsync(c->H, 0, s->hbytes, c->stream));
    hipLaunchKernelGGL(adj_init_pairs_kernel, dim3((c->V + 255) / 256), dim3(256), 0, c->stream, c->H, s->initbuf,
                       c->NP, c->V, c->N, 0);
    LAUNCHCHK(c);
  }
  c->st = state_of_step(2 + j * C, c->N);
  int rc;
  if ((rc = upload_state(c))) return rc;
  int nfree = 0;
  for (int b : s->hseg) nfree += b < 0 ? 1 : 0;
  for (int cur = j; cur < k;) {
    const int n = k - cur + 1;  // segments cur .. k still to be reversed
    const int bm = n >= 2 ? s->rvF[(size_t)n * s->rv_w + nfree] : 0;
    const int m = bm > 0 ? cur + bm : k;
    if ((rc = advance_steps(c, 3 + cur * C, 2 + m * C))) return rc;
    cur = m;
    if (cur < k) {
      if ((rc = store_snapshot(c, cur))) return rc;
      --nfree;
    }
  }
  return KB_OK;
}

// restore the checkpoint before step `first` and recompute steps
// first .. last, keeping their dropped columns on the tape
int recompute_segment(kb_ctx* c, int first, int last) {
  AdjState* s = c->adj;
  const int k = (first - 3) / s->C;
  int rc;
  if (s->hwin >= 0) {
    if ((rc = host_window(c, k))) return rc;
  } else {
    HIPCHK(c, hipMemcpyAsync(c->H, s->ckpt[k], s->hbytes, hipMemcpyDeviceToDevice, c->stream));
    c->st = state_of_step(first - 1, c->N);
    if ((rc = upload_state(c))) return rc;
  }
  for (int t = first; t <= last; ++t) {
    advance_state(c->st, c->N);
    if ((rc = save_column(c, t))) return rc;
    if ((rc = enqueue_step(c))) return rc;
    s->recomputed += 1;
  }
  return KB_OK;
}

}  // namespace

int adj_before_step(kb_ctx* c) {
  AdjState* s = c->adj;
  const int t = c->st.t;
  if (t > s->max_steps)
    return seterr(c, KB_ERR_ARG, "adjoint recording: step %d beyond max_steps %d", t, s->max_steps);
  return save_column(c, t);
}

int adj_after_step(kb_ctx* c) {
  AdjState* s = c->adj;
  const int t = c->st.t;
  if (s->C > 0 && (t - 2) % s->C == 0) {
    const int k = (t - 2) / s->C;
    if (s->hwin >= 0) {
      if (std::find(s->plan.begin(), s->plan.end(), k) != s->plan.end()) return store_snapshot(c, k);
    } else if ((size_t)k < s->ckpt.size()) {
      HIPCHK(c, hipMemcpyAsync(s->ckpt[k], c->H, s->hbytes, hipMemcpyDeviceToDevice, c->stream));
    }
  }
  return KB_OK;
}

void adj_free(kb_ctx* c) {
  if (c && c->adj) {
    c->bytes -= c->adj->bytes;
    free_state(c->adj);
    c->adj = nullptr;
  }
}

}  // namespace qftkb_host
#pragma GCC visibility pop

namespace {

// host_windows < 0: checkpoints (if any) in device memory
int adj_record(kb_ctx* c, int checkpoint_every, int max_steps, int host_windows) {
  if (!c || checkpoint_every < -1 || max_steps < 2) return KB_ERR_ARG;
  if (!c->initialised || c->st.tmax != 2)
    return seterr(c, KB_ERR_STATE, "adjoint recording starts right after initialisation (window [1,2])");
  if (c->in_ensemble) return seterr(c, KB_ERR_STATE, "adjoint: not available for ensemble members");
  if (c->sym) return seterr(c, KB_ERR_STATE, "adjoint: not available with cubic-orbit storage");
  if (c->p.pexp != KB_PEXP_LAMBDA_LO && !sunset_expansion(c->p.pexp))
    return seterr(c, KB_ERR_ARG, "adjoint: lambda expansions only");
  {
    // per function, shared by every context: the largest supported window
    const size_t lds = memint_lds(640);
    hipError_t e = hipFuncSetAttribute((const void*)adj_reduce_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess)
      e = hipFuncSetAttribute((const void*)adj_apply_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)apply_lds());
    if (e != hipSuccess)
      return seterr(c, KB_ERR_ARG, "adjoint: %zu bytes of LDS rejected: %s", lds, hipGetErrorString(e));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  adj_free(c);
  drop_graph(c);
  AdjState* s = new (std::nothrow) AdjState();
  if (!s) return KB_ERR_OOM;
  s->C = checkpoint_every < 0 ? 0 : checkpoint_every;
  s->packed = checkpoint_every < 0;
  s->progress = std::getenv("QFTKB_PROGRESS") != nullptr;
  s->max_steps = max_steps;
  const size_t hcount = (size_t)c->NB * (size_t)c->NP * MB * 2;
  s->hbytes = hcount * sizeof(double);
  const size_t ring = (size_t)c->NB * c->N * MB * 2;
  s->col_doubles = ring;
  if (s->packed) {
    // the packed tape keeps one value per orbit: the state must be carried
    // exactly by its orbits (kL2 and the initial rows orbit-constant)
    std::vector<int> orbit_of, rep_site;
    cubic_orbits(c->p.Nx, c->p.sdim, orbit_of, rep_site);
    std::vector<double> lat((size_t)c->V);
    auto symmetric = [&](const std::vector<double>& v) {
      double scale = 0.0, dev = 0.0;
      for (int p = 0; p < c->V; ++p) {
        scale = std::max(scale, std::fabs(v[p]));
        dev = std::max(dev, std::fabs(v[p] - v[rep_site[orbit_of[p]]]));
      }
      return dev <= 1e-12 * scale;
    };
    bool ok = true;
    if (hipMemcpy(lat.data(), c->kL2, sizeof(double) * c->V, hipMemcpyDeviceToHost) != hipSuccess) ok = false;
    ok = ok && symmetric(lat);
    const int rows[3][2] = {{1, 1}, {2, 1}, {2, 2}};
    for (int k = 0; k < 3 && ok; ++k) {
      hipLaunchKernelGGL(gather_pair, dim3((c->V + 255) / 256), dim3(256), 0, c->stream, c->H, c->NP, c->V,
                         pair_off((rows[k][0] - 1) % c->N, (rows[k][1] - 1) % c->N, c->N), 0, 1.0, c->vbuf);
      ok = hipMemcpyAsync(lat.data(), c->vbuf, sizeof(double) * c->V, hipMemcpyDeviceToHost, c->stream) ==
               hipSuccess &&
           hipStreamSynchronize(c->stream) == hipSuccess && symmetric(lat);
    }
    if (!ok) {
      delete s;
      return seterr(c, KB_ERR_ARG, "adjoint: the packed tape needs a cubic-symmetric state (kL2 and initial rows)");
    }
    s->norb = (int)rep_site.size();
    s->col_doubles = (size_t)s->norb * c->N * 2;
    if (hipMalloc((void**)&s->orb_rep, sizeof(int) * s->norb) != hipSuccess ||
        hipMalloc((void**)&s->site_orb, sizeof(int) * c->V) != hipSuccess ||
        hipMemcpy(s->orb_rep, rep_site.data(), sizeof(int) * s->norb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->site_orb, orbit_of.data(), sizeof(int) * c->V, hipMemcpyHostToDevice) != hipSuccess) {
      free_state(s);
      return seterr(c, KB_ERR_OOM, "adjoint: orbit maps");
    }
  }
  if (s->C > 0)
    s->tape_cols = s->C;
  else
    s->tape_cols = max_steps > c->N ? max_steps - c->N : 0;
  int rc = KB_OK;
  auto fail = [&](int code) {
    free_state(s);
    return code;
  };
  for (int j = 0; j < AK; ++j)
    if ((rc = amalloc(c, s, (void**)&s->Sring[j], ring * 8)) || (rc = amalloc(c, s, (void**)&s->Rring[j], ring * 8)))
      return fail(rc);
  if ((rc = amalloc(c, s, (void**)&s->Hbar, s->hbytes)) || (rc = amalloc(c, s, (void**)&s->Sbar, ring * 8)) ||
      (rc = amalloc(c, s, (void**)&s->X, (size_t)c->N * c->V * sizeof(double2))) ||
      (rc = amalloc(c, s, (void**)&s->P, 2 * (size_t)c->NB * 8)) ||
      (rc = amalloc(c, s, (void**)&s->LP, (size_t)c->NB * 8)) || (rc = amalloc(c, s, (void**)&s->m2r, 4 * 8)) ||
      (rc = amalloc(c, s, (void**)&s->Mb, 8)) || (rc = amalloc(c, s, (void**)&s->out, 4 * (size_t)c->V * 8)) ||
      (rc = amalloc(c, s, (void**)&s->tape, (size_t)s->tape_cols * s->col_doubles * 8)))
    return fail(rc);
  if (s->C > 0 && host_windows >= 0) {
    // windows in pinned host memory, the initial one as its three rows
    s->hwin = host_windows;
    // the windows are touched only when first written: refuse up front what
    // the machine's free physical memory cannot hold
    const double avail = (double)sysconf(_SC_AVPHYS_PAGES) * (double)sysconf(_SC_PAGESIZE);
    if ((double)host_windows * (double)s->hbytes > 0.9 * avail)
      return fail(seterr(c, KB_ERR_OOM, "adjoint: %d host windows of %zu bytes exceed 90%% of the %.0f free bytes",
                         host_windows, s->hbytes, avail));
    for (int i = 0; i < host_windows; ++i) {
      char* p = static_cast<char*>(std::malloc(s->hbytes));
      if (!p)
        return fail(seterr(c, KB_ERR_OOM, "adjoint: %d host windows of %zu bytes do not fit in host memory",
                           host_windows, s->hbytes));
      s->hsnap.push_back(p);
      s->hseg.push_back(-1);
      s->host_bytes += s->hbytes;
    }
    for (int k = 0; k < 2 && host_windows > 0; ++k)
      if (hipHostMalloc((void**)&s->stage[k], STAGE_BYTES, hipHostMallocDefault) != hipSuccess ||
          hipEventCreateWithFlags(&s->stage_ev[k], hipEventDisableTiming) != hipSuccess)
        return fail(seterr(c, KB_ERR_OOM, "adjoint: pinned staging buffers"));
    const int nseg = (max_steps - 2 + s->C - 1) / s->C;
    revolve_tables(s, nseg, host_windows);
    s->plan = host_plan(s, nseg, host_windows);
    if ((rc = amalloc(c, s, (void**)&s->initbuf, 6 * (size_t)c->V * sizeof(double)))) return fail(rc);
    hipLaunchKernelGGL(adj_init_pairs_kernel, dim3((c->V + 255) / 256), dim3(256), 0, c->stream, c->H, s->initbuf,
                       c->NP, c->V, c->N, 1);
    if (hipGetLastError() != hipSuccess) return fail(seterr(c, KB_ERR_HIP, "adjoint: initial rows not saved"));
  } else if (s->C > 0) {
    const int nck = (max_steps - 2) / s->C + 1;  // windows at t = 2, 2 + C, ...
    for (int k = 0; k < nck; ++k) {
      double* p = nullptr;
      if ((rc = amalloc(c, s, (void**)&p, s->hbytes))) return fail(rc);
      s->ckpt.push_back(p);
    }
    if (hipMemcpyAsync(s->ckpt[0], c->H, s->hbytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
      return fail(seterr(c, KB_ERR_HIP, "adjoint: checkpoint copy failed"));
  }
  bool ok = hipMemsetAsync(s->Hbar, 0, s->hbytes, c->stream) == hipSuccess &&
            hipMemsetAsync(s->Sbar, 0, ring * 8, c->stream) == hipSuccess;
  // the coefficient rings are read for every slot and lane (padding lanes and
  // slots outside a window are selected away): keep them finite
  for (int j = 0; j < AK && ok; ++j)
    ok = hipMemsetAsync(s->Sring[j], 0, ring * 8, c->stream) == hipSuccess &&
         hipMemsetAsync(s->Rring[j], 0, ring * 8, c->stream) == hipSuccess;
  if (!ok || hipStreamSynchronize(c->stream) != hipSuccess)
    return fail(seterr(c, KB_ERR_HIP, "adjoint: initialisation failed"));
  c->adj = s;
  c->bytes += s->bytes;
  return KB_OK;
}

}  // namespace

extern "C" {

int kb_adj_record(kb_ctx* c, int checkpoint_every, int max_steps) {
  return adj_record(c, checkpoint_every, max_steps, -1);
}

int kb_adj_record_host(kb_ctx* c, int checkpoint_every, int max_steps, int host_windows) {
  if (!c || checkpoint_every < 1 || host_windows < 0) return KB_ERR_ARG;
  return adj_record(c, checkpoint_every, max_steps, host_windows);
}

int kb_adj_stop(kb_ctx* c) {
  if (!c) return KB_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  adj_free(c);
  return KB_OK;
}

int kb_adj_info(kb_ctx* c, long long* device_bytes, long long* recomputed_steps) {
  if (!c) return KB_ERR_ARG;
  if (!c->adj) return seterr(c, KB_ERR_STATE, "adjoint: not recording");
  if (device_bytes) *device_bytes = (long long)c->adj->bytes;
  if (recomputed_steps) *recomputed_steps = c->adj->recomputed;
  return KB_OK;
}

int kb_adj_debug_stop(kb_ctx* c, int t_stop) {
  if (!c || t_stop < 3) return KB_ERR_ARG;
  if (!c->adj) return seterr(c, KB_ERR_STATE, "adjoint: not recording");
  c->adj->stop_after = t_stop;
  return KB_OK;
}

int kb_adj_get_entry(kb_ctx* c, int which, int i, int j, double* out) {
  if (!c || !out || (which != 0 && which != 1)) return KB_ERR_ARG;
  if (!c->adj) return seterr(c, KB_ERR_STATE, "adjoint: not recording");
  if (i < c->st.tmin || i > c->st.tmax || j < c->st.tmin || j > c->st.tmax)
    return seterr(c, KB_ERR_WINDOW, "adjoint get: (%d,%d) outside window [%d,%d]", i, j, c->st.tmin, c->st.tmax);
  const int si = (i - 1) % c->N, sj = (j - 1) % c->N;
  const double sg = which ? (si >= sj ? 1.0 : -1.0) : 1.0;
  hipLaunchKernelGGL(gather_pair, dim3((c->V + 255) / 256), dim3(256), 0, c->stream, c->adj->Hbar, c->NP, c->V,
                     pair_off(si, sj, c->N), which, sg, c->vbuf);
  LAUNCHCHK(c);
  return download(c, out, c->vbuf, c->V);
}

int kb_adj_gradient(kb_ctx* c, const double* wgt, double* dL_dlambda, double* dL_dn0, double* dL_dF0) {
  if (!c || !wgt) return KB_ERR_ARG;
  AdjState* s = c->adj;
  if (!s) return seterr(c, KB_ERR_STATE, "adjoint: kb_adj_record was not called");
  if (dL_dn0 && !c->init_particles)
    return seterr(c, KB_ERR_STATE, "adjoint: dL/dn0 needs initialisation by kb_init_particles");
  const int T = c->st.tmax;
  const int N = c->N;
  if (T < 2 || T > s->max_steps) return seterr(c, KB_ERR_STATE, "adjoint: window end %d not recorded", T);
  int rc;
  s->recomputed = 0;
  HIPCHK(c, hipMemsetAsync(s->Hbar, 0, s->hbytes, c->stream));
  HIPCHK(c, hipMemsetAsync(s->P, 0, 2 * (size_t)c->NB * 8, c->stream));
  HIPCHK(c, hipMemsetAsync(s->LP, 0, (size_t)c->NB * 8, c->stream));
  HIPCHK(c, hipMemsetAsync(s->m2r, 0, 4 * 8, c->stream));
  if ((rc = upload(c, c->vbuf, wgt, c->V))) return rc;
  const int blocks = (c->V + 255) / 256;
  hipLaunchKernelGGL(adj_seed_kernel, dim3(blocks), dim3(256), 0, c->stream, c->H, s->Hbar, c->vbuf, c->NP, c->V,
                     N, (T - 1) % N, (T - 2) % N, c->p.dt);
  LAUNCHCHK(c);
  // the tape holds the last segment's columns already (or every column)
  int have_from = s->C > 0 ? seg_first(T, s->C) : 3;
  const int stop = s->stop_after;
  s->stop_after = 0;
  int npend = 0;  // reverse steps whose entry updates are deferred
  for (int t = T; t >= 3; --t) {
    if (t < have_from) {
      const int first = seg_first(t, s->C);
      if ((rc = recompute_segment(c, first, t))) return rc;
      have_from = first;
      if (s->progress) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        std::fprintf(stderr, "[qftkb] adjoint: reversing steps %d .. %d, %lld steps recomputed so far\n", t, first,
                     s->recomputed);
        std::fflush(stderr);
      }
    }
    const bool apply = npend + 1 == AK || t == 3 || t == stop;
    if ((rc = reverse_step(c, t, T, npend, apply))) return rc;
    npend = apply ? 0 : npend + 1;
    if (t == stop) {
      // left mid-sweep for inspection: the window of step t-1
      c->st = state_of_step(t - 1, N);
      if ((rc = upload_state(c))) return rc;
      HIPCHK(c, hipStreamSynchronize(c->stream));
      return KB_OK;
    }
  }
  const double h = (double)(c->p.ONgroup + 2) / (6.0 * (double)c->p.ONgroup);
  const bool pending = T >= 3;
  if (pending) {
    hipLaunchKernelGGL(adj_pending_kernel, dim3(1), dim3(1024), 0, c->stream, s->P, c->NB, c->p.dt,
                       h / (double)c->V, s->m2r, s->LP, s->Mb);
    LAUNCHCHK(c);
  }
  hipLaunchKernelGGL(adj_final_kernel, dim3(blocks), dim3(256), 0, c->stream, s->Hbar, c->NP, c->V, N, c->kL2,
                     c->initmass, c->init_particles ? 1 : 0, c->p.dt, c->p.Lambda * h / (double)c->V,
                     pending ? s->Mb : nullptr, s->out);
  LAUNCHCHK(c);
  // the window is back at the initial state [1, 2]
  c->st = state_of_step(2, N);
  if ((rc = upload_state(c))) return rc;
  std::vector<double> lp((size_t)c->NB);
  if ((rc = download(c, lp.data(), s->LP, lp.size()))) return rc;
  double lam = 0.0;
  for (double v : lp) lam += v;
  if (dL_dlambda) *dL_dlambda = lam;
  if (dL_dF0 && (rc = download(c, dL_dF0, s->out, 3 * (size_t)c->V))) return rc;
  if (dL_dn0 && (rc = download(c, dL_dn0, s->out + 3 * (size_t)c->V, c->V))) return rc;
  // the initial window is checkpoint 0 again; a new forward run may be recorded
  for (int& b : s->hseg) b = -1;
  return KB_OK;
}

}  // extern "C"
