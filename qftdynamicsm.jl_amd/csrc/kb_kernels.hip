// kb_kernels.hip -- gfx950 kernels of the 2PI Kadanoff-Baym step.
//
//   sigma_pass     self-energy: batched pencil FFTs staged in LDS with the
//                  sunset products fused into the last forward pass
//                  (TwoPIScalar_Evolution.jl:16-31)
//   hartree        Hartree mass^2 reduction (TwoPIScalar_Initialization.jl:
//                  207-218)
//   memint_step    memory integrals + leapfrog + diagonal of one KB step,
//                  one workgroup per mode block, history read once
//                  (TwoPIScalar_Evolution.jl:766-782, 846-854, 890-915,
//                  951-974)
//   measure        equal-time observables and energies
//                  (TwoPIScalar_Measurement.jl:31-99, 135-174)
// plus small gather/scatter kernels for the host API.
#include "kb_kernels.h"

#include <math.h>

namespace qftkb {

// ---------------------------------------------------------------------------
// LDS FFT (Stockham, mixed radix)
// ---------------------------------------------------------------------------
__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// exp(-+ 2 pi i k / n) from the table, conjugated for the inverse
__device__ inline double2 twid(const double2* __restrict__ tw, int k, bool inv) {
  double2 w = tw[k];
  if (inv) w.y = -w.y;
  return w;
}

template <int R>
__device__ inline void small_dft(double2 (&v)[R], const double2* __restrict__ tw, int n, bool inv) {
  if constexpr (R == 2) {
    double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    // forward: multiply by -i for the odd quarter; inverse: +i
    double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    double2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
    double2 jb1 = inv ? make_double2(-b1.y, b1.x) : make_double2(b1.y, -b1.x);
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, jb1);
    v[3] = csub(a1, jb1);
  } else {
    double2 o[R];
    const int m = n / R;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      double2 s = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) s = cadd(s, cmul(v[r], twid(tw, ((q * r) % R) * m, inv)));
      o[q] = s;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = o[q];
  }
}

template <int R>
__device__ inline void stockham_stage(const double2* __restrict__ src, double2* __restrict__ dst, int n,
                                      int P, int Ns, const double2* __restrict__ tw, bool inv) {
  const int m = n / R;
  const int step = n / (Ns * R);
  for (int idx = threadIdx.x; idx < P * m; idx += blockDim.x) {
    const int k = idx / m;
    const int j = idx - k * m;
    const double2* s = src + k * n;
    double2* d = dst + k * n;
    const int kk = j % Ns;
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      double2 x = s[j + r * m];
      v[r] = r == 0 ? x : cmul(x, twid(tw, (r * kk * step) % n, inv));
    }
    small_dft<R>(v, tw, n, inv);
    const int base = (j / Ns) * Ns * R + kk;
#pragma unroll
    for (int q = 0; q < R; ++q) d[base + q * Ns] = v[q];
  }
}

// radix > 8: direct sum per output element
__device__ inline void direct_stage(const double2* __restrict__ src, double2* __restrict__ dst, int n, int P,
                                    int Ns, int R, const double2* __restrict__ tw, bool inv) {
  const int m = n / R;
  const int step = n / (Ns * R);
  for (int idx = threadIdx.x; idx < P * n; idx += blockDim.x) {
    const int k = idx / n;
    const int rem = idx - k * n;
    const int j = rem / R;
    const int q = rem - j * R;
    const double2* s = src + k * n;
    const int kk = j % Ns;
    double2 acc = make_double2(0.0, 0.0);
    for (int r = 0; r < R; ++r) {
      double2 x = s[j + r * m];
      if (r) x = cmul(x, twid(tw, (int)(((long)r * kk * step) % n), inv));
      acc = cadd(acc, cmul(x, twid(tw, ((q * r) % R) * m, inv)));
    }
    dst[k * n + (j / Ns) * Ns * R + kk + q * Ns] = acc;
  }
}

// Transforms P pencils of length n held in a (ping-pong with b).  Returns
// the buffer holding the result.  Ends with a barrier.
__device__ double2* lds_fft(double2* a, double2* b, int n, int P, const FFTPlan& pl,
                            const double2* __restrict__ tw, bool inv) {
  int Ns = 1;
  for (int s = 0; s < pl.nf; ++s) {
    const int R = pl.fac[s];
    switch (R) {
      case 2: stockham_stage<2>(a, b, n, P, Ns, tw, inv); break;
      case 3: stockham_stage<3>(a, b, n, P, Ns, tw, inv); break;
      case 4: stockham_stage<4>(a, b, n, P, Ns, tw, inv); break;
      case 5: stockham_stage<5>(a, b, n, P, Ns, tw, inv); break;
      case 7: stockham_stage<7>(a, b, n, P, Ns, tw, inv); break;
      default: direct_stage(a, b, n, P, Ns, R, tw, inv); break;
    }
    __syncthreads();
    double2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// ---------------------------------------------------------------------------
// Self-energy pass.  One workgroup: row q (t' = tmin + q) x P pencils along
// `axis`.  See SigmaArgs for the flags.
// ---------------------------------------------------------------------------
__device__ inline int site_of(const SigmaArgs& a, int pencil, int j) {
  // dims (nx, ny, nz); pencil enumerates the two non-transform coordinates
  if (a.axis == 0) {
    const int iy = pencil % a.ny, iz = pencil / a.ny;
    return j + a.nx * (iy + a.ny * iz);
  } else if (a.axis == 1) {
    const int ix = pencil % a.nx, iz = pencil / a.nx;
    return ix + a.nx * (j + a.ny * iz);
  } else {
    const int ix = pencil % a.nx, iy = pencil / a.nx;
    return ix + a.nx * (iy + a.ny * j);
  }
}

__device__ inline int neg_site(const SigmaArgs& a, int s) {
  const int ix = s % a.nx;
  const int iy = (s / a.nx) % a.ny;
  const int iz = s / (a.nx * a.ny);
  const int jx = (a.nx - ix) % a.nx, jy = (a.ny - iy) % a.ny, jz = (a.nz - iz) % a.nz;
  return jx + a.nx * (jy + a.ny * jz);
}

__global__ __launch_bounds__(256) void sigma_pass(SigmaArgs a) {
  extern __shared__ double2 lds[];
  const int n = a.n;
  const int P = a.P;
  double2* A = lds;
  double2* B = lds + P * n;
  const int q = blockIdx.y;
  if (q >= a.st->L) return;  // rows t' = tmin + q, q < L, are live
  const int pen0 = blockIdx.x * P;
  const int tp = a.st->tmin + q;
  const int s_tm1 = a.st->s_tm1;
  const int slot = (tp - 1) % a.N;
  const size_t rowoff = (size_t)q * a.V;
  const bool contiguous = a.axis == 0;

  // ---- load ----
  for (int e = threadIdx.x; e < P * n; e += blockDim.x) {
    int k, j;
    if (contiguous) {
      k = e / n;
      j = e - k * n;
    } else {
      j = e / P;
      k = e - j * P;
    }
    const int pencil = pen0 + k;
    double2 z = make_double2(0.0, 0.0);
    if (pencil < a.npencils) {
      const int s = site_of(a, pencil, j);
      if (a.flags & SIGMA_GATHER) {
        // physical F(t-1,t') and rho(t-1,t'); rho sign from the slot order
        const long pr = pair_off(s_tm1, slot, a.N);
        const double sg = s_tm1 >= slot ? 1.0 : -1.0;
        const int s2 = neg_site(a, s);
        const double* h1 = a.H + hoff(s / MB, pr, s % MB, a.NP);
        const double* h2 = a.H + hoff(s2 / MB, pr, s2 % MB, a.NP);
        // real(rfft(X)) only sees the inversion-even part of X
        z.x = 0.5 * (h1[0] + h2[0]);
        z.y = 0.5 * (h1[1] + h2[1]) * sg;
      } else {
        z = a.W[rowoff + s];
      }
    }
    A[k * n + j] = z;
  }
  __syncthreads();

  // ---- transform(s) ----
  double2* res = lds_fft(A, B, n, P, a.plan, a.tw, (a.flags & SIGMA_INVERSE) != 0);
  if (a.flags & SIGMA_SUNSET) {
    double2* other = res == A ? B : A;
    // quantum sunset; the classical one drops the rho^2 terms
    const double qF = a.classical ? 0.0 : 0.75, qr = a.classical ? 0.0 : 1.0 / 12.0;
    for (int e = threadIdx.x; e < P * n; e += blockDim.x) {
      const double2 z = res[e];
      const double Fx = z.x, rx = z.y;
      const double FF = Fx * Fx, rr = rx * rx;
      res[e] = make_double2(Fx * (FF - qF * rr), rx * (FF - qr * rr));
    }
    __syncthreads();
    res = lds_fft(res, other, n, P, a.plan, a.tw, true);
  }

  // ---- store ----
  for (int e = threadIdx.x; e < P * n; e += blockDim.x) {
    int k, j;
    if (contiguous) {
      k = e / n;
      j = e - k * n;
    } else {
      j = e / P;
      k = e - j * P;
    }
    const int pencil = pen0 + k;
    if (pencil >= a.npencils) continue;
    const int s = site_of(a, pencil, j);
    const double2 z = res[k * n + j];
    if (a.flags & SIGMA_SCATTER) {
      double* o = a.S + soff(s / MB, slot, s % MB, a.N);
      o[0] = a.preF * z.x;
      o[1] = a.preR * z.y;
    } else {
      a.W[rowoff + s] = z;
    }
  }
}

// ---------------------------------------------------------------------------
// Hartree mass: m2 = Mass2 + H ivol sum_p F(t-1,t-1;p)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void hartree_kernel(const double* __restrict__ H, long NP, int V, long pair,
                                                       double hfac, double Mass2, const double* __restrict__ wgt,
                                                       double* __restrict__ out) {
  __shared__ double red[1024];
  double acc = 0.0;
  for (int p = threadIdx.x; p < V; p += blockDim.x)
    acc += (wgt ? wgt[p] : 1.0) * H[hoff(p / MB, pair, p % MB, NP)];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double h = red[0];
    h *= hfac;
    h += Mass2;
    out[0] = h;
  }
}

// One workgroup per HARTREE_SITES sites, HARTREE_LOADS loads in flight per
// thread (the diagonal entries lie one per 128-byte line of every site
// block, so a single workgroup is latency-bound).  Each workgroup writes its
// partial sum; the last one to finish (an agent-scope counter: the
// workgroups sit on different XCDs) adds the partials in workgroup order, so
// the result does not depend on which one that is, and resets the counter.
__global__ __launch_bounds__(HARTREE_THREADS) void hartree_step_kernel(const double* __restrict__ H, long NP, int V,
                                                                       int N, const StepState* __restrict__ st,
                                                                       double hfac, double Mass2,
                                                                       const double* __restrict__ wgt,
                                                                       double* __restrict__ out, double* part,
                                                                       double* __restrict__ sum_out) {
  __shared__ double red[HARTREE_THREADS / 64];
  __shared__ int is_last;
  const long pair = pair_off(st->s_tm1, st->s_tm1, N);  // F(t-1,t-1), Evolution.jl:945
  const int G = gridDim.x;
  double v[HARTREE_LOADS];
#pragma unroll
  for (int u = 0; u < HARTREE_LOADS; ++u) {
    const int p = blockIdx.x * HARTREE_SITES + u * HARTREE_THREADS + threadIdx.x;
    v[u] = p < V ? (wgt ? wgt[p] : 1.0) * H[hoff(p / MB, pair, p % MB, NP)] : 0.0;
  }
  double acc = 0.0;
#pragma unroll
  for (int u = 0; u < HARTREE_LOADS; ++u) acc += v[u];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  unsigned* count = reinterpret_cast<unsigned*>(part + G);
  if (threadIdx.x == 0) {
    double b = 0.0;
    for (int w = 0; w < HARTREE_THREADS / 64; ++w) b += red[w];
    part[blockIdx.x] = b;
    is_last = __hip_atomic_fetch_add(count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == (unsigned)G - 1;
  }
  __syncthreads();
  if (!is_last) return;
  // the partials in a fixed order: thread k takes k, k + T, ..., then the
  // same tree as above
  double h = 0.0;
  for (int g = threadIdx.x; g < G; g += HARTREE_THREADS) h += part[g];
  for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = h;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int w = 0; w < HARTREE_THREADS / 64; ++w) tot += red[w];
    out[0] = tot * hfac + Mass2;
    if (sum_out) sum_out[0] = tot;
    __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void advance_kernel(StepState* const* states, const int* Ns, int nruns) {
  const int r = threadIdx.x;
  if (r < nruns) {
    StepState s = *states[r];
    advance_state(s, Ns[r]);
    *states[r] = s;
  }
}

// ---------------------------------------------------------------------------
// Memory integrals + leapfrog + diagonal, one workgroup per mode block.
//
// With rho antisymmetric and F symmetric, the reference's
//   R_r(t') = -sum_{z=t'+1}^{t-2} Sr(z) r(z,t')                    (Evolution.jl:846-854)
//   R_F(t') = -sum_{z=tmin}^{t-2} w_z Sr(z) F(z,t')
//             +sum_{z=tmin}^{t'-1} w_z SF(z) r(z,t')                (Evolution.jl:766-782)
// (w = 1/2 at tmin, 1 otherwise; S(z) = Sigma(t-1,z)) make every stored
// entry (a,b), a>b, feed two output rows: with u = m w Sr, v = w SF,
// ms = m Sr (m = 0 at z = t-1):
//   R_F(a) += -u(b) F - v(b) r,  R_F(b) += -u(a) F,  R_r(b) += -ms(a) r,
// and the diagonal (a,a) feeds R_F(a) += -u(a) F.  The kernel walks the
// slot triangle once: every 8-lane worker streams chunks of rows of one
// 4-column block (contiguous in HBM) with two row batches in flight,
// keeps the 4 column sums in registers and adds row and column sums into
// per-slot LDS accumulators with ds_add_f64.
// In slot space, pair (hi, lo) has time(hi) < time(lo) exactly when
// lo < s_min <= hi; those pairs swap the roles of a and b.
// ---------------------------------------------------------------------------
// Tile shape (template parameters, picked per lattice and window by the
// autotuner): ROWS slot rows per work item (one chunk of a column block),
// B rows per batch, two batches in flight.

__device__ __forceinline__ void slot_weights(int x, int smin, int N, int L, double& w, double& m) {
  const int rel = window_pos(x, smin, N);
  w = rel == 0 ? 0.5 : 1.0;
  m = rel == L - 1 ? 0.0 : 1.0;
}

// History lines are streamed once per step: load them non-temporally so
// they do not evict the self-energy lines that are re-read from L2.
#ifndef KB_NT_LOADS
#define KB_NT_LOADS 1
#endif
// Loads go through global-address-space pointers: the pointers come from the
// argument struct, so plain dereferences compile to flat loads, whose waits
// also drain the LDS queue and every load in flight (s_waitcnt vmcnt(0)
// lgkmcnt(0)); global loads let a batch wait only for its own lines.
typedef double kb_d2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const kb_d2v kb_gd2v;
__device__ __forceinline__ double2 ld_stream(const double* p) {
#if KB_NT_LOADS
  const kb_d2v v = __builtin_nontemporal_load((kb_gd2v*)(p));
#else
  const kb_d2v v = *(kb_gd2v*)(p);
#endif
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ double2 ld_global(const void* p) {
  const kb_d2v v = *(kb_gd2v*)(p);
  return make_double2(v.x, v.y);
}

template <int B>
struct RowBatch {
  double2 e[B][CB];  // (F, r) of slots (row, c0 + jc)
  double2 fs[B];     // (SigF, Sigr) of the row slot
};

// rows r .. r+MI_B-1 of the column block starting at storage slot cbase;
// rows past rlast are clamped to rlast (re-read, not used)
template <int B>
__device__ __forceinline__ void load_batch(RowBatch<B>& b, const double* __restrict__ Hb,
                                           const double2* __restrict__ Sb, long cbase, int c0, int r, int rlast) {
#pragma unroll
  for (int h = 0; h < B; ++h) {
    const int rr = min(r + h, rlast);
    const long pb = cbase + (long)(rr - c0) * CB;
#pragma unroll
    for (int jc = 0; jc < CB; ++jc) b.e[h][jc] = ld_stream(Hb + (size_t)(pb + jc) * MB * 2);
    b.fs[h] = ld_global(Sb + (size_t)rr * MB);
  }
}

struct ColState {
  double u[CB], v[CB], ms[CB], cF[CB], cR[CB];
  unsigned phi;  // bit jc: column slot lies before s_min (its time is after the wrap)
};

// Pair (r, c), r > c in slot space.  Time order is reversed (time(r) <
// time(c)) exactly when r >= s_min > c; `neg` carries that per column.
__device__ __forceinline__ void pair_update(double Fv, double Pv, bool n, double uc, double vc, double msc,
                                            double ur, double vr, double msr, double& rowF, double& rowR,
                                            double& cF, double& cR) {
  rowF -= uc * Fv + (n ? 0.0 : vc) * Pv;
  rowR += (n ? msc : 0.0) * Pv;
  cF += (n ? vr : 0.0) * Pv - ur * Fv;
  cR -= (n ? 0.0 : msr) * Pv;
}

template <int B>
__device__ __forceinline__ void process_batch(const RowBatch<B>& b, int r, int r1, int dead, int smin, int N, int L,
                                              ColState& cs, double* accF, double* accR, int lane) {
#pragma unroll
  for (int h = 0; h < B; ++h) {
    const int rr = r + h;
    if (rr < r1) {
      const bool okr = rr != dead;
      double w, m;
      slot_weights(rr, smin, N, L, w, m);
      const double ur = okr ? m * w * b.fs[h].y : 0.0;
      const double vr = okr ? w * b.fs[h].x : 0.0;
      const double msr = okr ? m * b.fs[h].y : 0.0;
      const unsigned neg = rr >= smin ? cs.phi : 0u;
      double rowF = 0.0, rowR = 0.0;
#pragma unroll
      for (int jc = 0; jc < CB; ++jc)
        pair_update(b.e[h][jc].x, b.e[h][jc].y, (neg >> jc) & 1u, cs.u[jc], cs.v[jc], cs.ms[jc], ur, vr, msr, rowF,
                    rowR, cs.cF[jc], cs.cR[jc]);
      atomicAdd(&accF[rr * MB + lane], rowF);
      atomicAdd(&accR[rr * MB + lane], rowR);
    }
  }
}

// Sweep of work items [ibeg, iend) of one site block: memory-integral row
// and column sums into the LDS accumulators (see above).  Workers start at
// ibeg + worker and take further items from the LDS counter *next_item.
template <int ROWS, int B>
__device__ __forceinline__ void memint_sweep(const MemintArgs& a, const StepState& ss, int mb, double* accF,
                                             double* accR, int* next_item, int ibeg, int iend) {
  const int N = a.N;
  const int tid = threadIdx.x;
  const int lane = tid & (MB - 1);
  const int worker = tid >> 3;
  const int Rmax = ss.Rmax;
  const int dead = ss.dead;
  const int smin = ss.s_min;
  const int L = ss.L;
  const double* __restrict__ Hb = a.H + hoff(mb, 0, 0, a.NP) + lane * 2;
  // self-energies of this block: (SigF, Sigr) per slot for this lane
  const double2* __restrict__ Sb = reinterpret_cast<const double2*>(a.S + soff(mb, 0, lane, N));
  // work items: (column block J, chunk of MI_ROWS rows), in storage order;
  // the first item of each worker is static, later ones are taken from an
  // LDS counter, so workers that finish early take more
  int item = ibeg + worker;
  while (item < iend) {
    const int2 jk = a.items[item];
    const int J = jk.x;
    const int c0 = J * CB;
    const int r0 = c0 + jk.y * ROWS;
    const int r1 = min(r0 + ROWS, Rmax);
    if (c0 < Rmax && r0 < r1) {
      const long cbase = colblock_offset(J, N);
      ColState cs;
      cs.phi = 0;
#pragma unroll
      for (int jc = 0; jc < CB; ++jc) {
        const int c = c0 + jc;
        const bool ok = c < Rmax && c != dead;
        const double2 fs = ld_global(Sb + (size_t)min(c, N - 1) * MB);
        double w, m;
        slot_weights(c, smin, N, L, w, m);
        cs.u[jc] = ok ? m * w * fs.y : 0.0;
        cs.v[jc] = ok ? w * fs.x : 0.0;
        cs.ms[jc] = ok ? m * fs.y : 0.0;
        cs.cF[jc] = cs.cR[jc] = 0.0;
        cs.phi |= (c < smin ? 1u : 0u) << jc;
      }
      int r = r0;
      // rows crossing the diagonal of the block: columns c0 .. r only
      for (; r < r1 && r < c0 + CB; ++r) {
        const bool okr = r != dead;
        const double2 fs = ld_global(Sb + (size_t)r * MB);
        double w, m;
        slot_weights(r, smin, N, L, w, m);
        const double ur = okr ? m * w * fs.y : 0.0;
        const double vr = okr ? w * fs.x : 0.0;
        const double msr = okr ? m * fs.y : 0.0;
        const unsigned neg = r >= smin ? cs.phi : 0u;
        const long pb = cbase + (long)(r - c0) * CB;
        double rowF = 0.0, rowR = 0.0;
        for (int jc = 0; jc < CB; ++jc) {
          const int c = c0 + jc;
          if (c > r) break;
          const double2 e = ld_global(Hb + (size_t)(pb + jc) * MB * 2);
          if (c == r)
            rowF -= ur * e.x;
          else
            pair_update(e.x, e.y, (neg >> jc) & 1u, cs.u[jc], cs.v[jc], cs.ms[jc], ur, vr, msr, rowF, rowR,
                        cs.cF[jc], cs.cR[jc]);
        }
        atomicAdd(&accF[r * MB + lane], rowF);
        atomicAdd(&accR[r * MB + lane], rowR);
      }
      // full rows: two batches of MI_B rows in flight
      if (r < r1) {
        const int rlast = r1 - 1;
        RowBatch<B> X, Y;
        load_batch(X, Hb, Sb, cbase, c0, r, rlast);
        while (true) {
          load_batch(Y, Hb, Sb, cbase, c0, r + B, rlast);
          process_batch(X, r, r1, dead, smin, N, L, cs, accF, accR, lane);
          r += B;
          if (r >= r1) break;
          load_batch(X, Hb, Sb, cbase, c0, r + B, rlast);
          process_batch(Y, r, r1, dead, smin, N, L, cs, accF, accR, lane);
          r += B;
          if (r >= r1) break;
        }
      }
#pragma unroll
      for (int jc = 0; jc < CB; ++jc) {
        const int c = c0 + jc;
        if (c < Rmax && c != dead) {
          atomicAdd(&accF[c * MB + lane], cs.cF[jc]);
          atomicAdd(&accR[c * MB + lane], cs.cR[jc]);
        }
      }
    }
    // next item: one LDS atomic per worker
    int nxt = 0;
    if (lane == 0) nxt = atomicAdd(next_item, 1);
    item = __shfl(nxt, tid & ~(MB - 1) & 63);
  }
}

// Leapfrog of row t and the diagonal from the accumulated memory integrals
// of one site block; red holds blockDim.x / 64 rows of MB doubles.
__device__ __forceinline__ void memint_leapfrog(const MemintArgs& a, const StepState& ss, int mb,
                                                const double* accF, const double* accR, double (*red)[MB],
                                                double (*diagF)[MB]) {
  const int N = a.N;
  const int tid = threadIdx.x;
  const int lane = tid & (MB - 1);
  const int worker = tid >> 3;
  const int nworkers = blockDim.x >> 3;
  const int Rmax = ss.Rmax;
  const int dead = ss.dead;
  const int smin = ss.s_min;
  const int L = ss.L;
  const double2* __restrict__ Sb = reinterpret_cast<const double2*>(a.S + soff(mb, 0, lane, N));
  // ---- leapfrog of row t (evolve_r!, evolve_F!: Evolution.jl:890-915) ----
  const int p = mb * MB + lane;
  const bool real_site = p < a.V;
  const double m2 = a.m2[0];
  const double dt = a.dt;
  const double dt2 = dt * dt, dt3 = dt * dt * dt;
  const double Om = (a.kL2[p] + m2) * dt2;  // setomega2values!, Evolution.jl:918-921
  double dsum = 0.0;
  double* __restrict__ Hw = a.H + hoff(mb, 0, 0, a.NP);
  constexpr int LU = 4;  // slots per worker in flight
  for (int x0 = worker * LU; x0 < Rmax; x0 += nworkers * LU) {
    double2 e1[LU], e2[LU], fsx[LU];
#pragma unroll
    for (int u = 0; u < LU; ++u) {
      const int x = min(x0 + u, Rmax - 1);
      e1[u] = ld_global(Hw + ((size_t)pair_off(ss.s_tm1, x, N) * MB + lane) * 2);
      e2[u] = ld_global(Hw + ((size_t)pair_off(ss.s_tm2, x, N) * MB + lane) * 2);
      fsx[u] = ld_global(Sb + (size_t)x * MB);
    }
#pragma unroll
    for (int u = 0; u < LU; ++u) {
      const int x = x0 + u;
      if (x >= Rmax || x == dead) continue;
      const int rel = window_pos(x, smin, N);
      const int tx = ss.tmin + rel;
      const double r1 = e1[u].y * (ss.s_tm1 >= x ? 1.0 : -1.0);
      const double r2 = e2[u].y * (ss.s_tm2 >= x ? 1.0 : -1.0);
      const double RHSF = accF[x * MB + lane] * dt3;
      const double Fn = (2.0 - Om) * e1[u].x - e2[u].x + RHSF;
      double rn;
      if (tx == ss.t - 1) {
        rn = dt;
      } else {
        const double RHSr = accR[x * MB + lane] * dt3;
        rn = (2.0 - Om) * r1 - r2 + RHSr;
      }
      if (real_site) {
        double2 o = make_double2(Fn, rn * (ss.s_t >= x ? 1.0 : -1.0));
        *reinterpret_cast<double2*>(Hw + ((size_t)pair_off(ss.s_t, x, N) * MB + lane) * 2) = o;
      }
      // terms of getRHS_F!(t,t) (Evolution.jl:973) over the fresh row
      if (a.integral) {
        const double w = rel == 0 ? 0.5 : 1.0;
        const double m = rel == L - 1 ? 0.0 : 1.0;
        dsum += -(m * w * fsx[u].y) * Fn - (w * fsx[u].x) * rn;
      }
      if (tx == ss.t - 1) diagF[0][lane] = Fn;
      if (tx == ss.t - 2) diagF[1][lane] = Fn;
    }
  }
  // reduce over the 8 workers of the wave (lane bits 3..5), then over waves
  dsum += __shfl_xor(dsum, 8);
  dsum += __shfl_xor(dsum, 16);
  dsum += __shfl_xor(dsum, 32);
  if ((tid & 63) < MB) red[tid >> 6][lane] = dsum;
  __syncthreads();
  if (tid < MB) {
    double RF = 0.0;
    for (int wv = 0; wv < (int)(blockDim.x >> 6); ++wv) RF += red[wv][tid];
    const double Ftt = (2.0 - Om) * diagF[0][tid] - diagF[1][tid] + RF * dt3;
    if (real_site) {
      *reinterpret_cast<double2*>(Hw + ((size_t)pair_off(ss.s_t, ss.s_t, N) * MB + tid) * 2) =
          make_double2(Ftt, 0.0);
    }
  }
}

template <int ROWS, int B, int TPB>
__global__ __launch_bounds__(TPB, 512 / TPB) void memint_step(const MemintArgs* __restrict__ runs,
                                                                 const int* __restrict__ wg_start, int nruns) {
  extern __shared__ double smem[];
  // which run this workgroup belongs to (runs are few; wg_start is uniform)
  int run = 0;
  while (run + 1 < nruns && (int)blockIdx.x >= wg_start[run + 1]) ++run;
  const MemintArgs a = runs[run];
  const StepState ss = *a.st;
  const int N = a.N;
  double* accF = smem;           // R_F per slot, lane
  double* accR = accF + N * MB;  // R_r
  __shared__ double red[TPB / 64][MB];
  __shared__ double diagF[2][MB];
  __shared__ int next_item;
  const int mb = blockIdx.x - wg_start[run];
  const int tid = threadIdx.x;
  for (int i = tid; i < N * MB; i += blockDim.x) {
    accF[i] = 0.0;
    accR[i] = 0.0;
  }
  if (tid == 0) next_item = blockDim.x >> 3;
  __syncthreads();
  if (a.integral) memint_sweep<ROWS, B>(a, ss, mb, accF, accR, &next_item, 0, a.nitems);
  __syncthreads();
  memint_leapfrog(a, ss, mb, accF, accR, red, diagF);
}

// Split form for windows with few site blocks (cubic-orbit storage): each
// site block's work items are cut into nsplit contiguous ranges, one
// workgroup each; memint_partial writes its accumulators to part, and
// memint_finish sums them and runs the leapfrog.
template <int ROWS, int B, int TPB>
__global__ __launch_bounds__(TPB, 512 / TPB) void memint_partial(const MemintArgs* __restrict__ runs, int nsplit,
                                                                    double* __restrict__ part) {
  extern __shared__ double smem[];
  const MemintArgs a = runs[0];
  const StepState ss = *a.st;
  const int N = a.N;
  double* accF = smem;
  double* accR = accF + N * MB;
  __shared__ int next_item;
  const int mb = blockIdx.x / nsplit;
  const int sp = blockIdx.x - mb * nsplit;
  const int ibeg = (int)((long)a.nitems * sp / nsplit);
  const int iend = (int)((long)a.nitems * (sp + 1) / nsplit);
  const int tid = threadIdx.x;
  for (int i = tid; i < N * MB; i += blockDim.x) {
    accF[i] = 0.0;
    accR[i] = 0.0;
  }
  if (tid == 0) next_item = ibeg + (blockDim.x >> 3);
  __syncthreads();
  if (a.integral) memint_sweep<ROWS, B>(a, ss, mb, accF, accR, &next_item, ibeg, iend);
  __syncthreads();
  double* out = part + (size_t)blockIdx.x * 2 * N * MB;
  for (int i = tid; i < 2 * N * MB; i += blockDim.x) out[i] = smem[i];
}

// (1024 threads: the sum over the splits' partials is latency-bound)
__global__ __launch_bounds__(MEMINT_FINISH_THREADS) void memint_finish(const MemintArgs* __restrict__ runs,
                                                                       int nsplit, const double* __restrict__ part) {
  extern __shared__ double smem[];
  const MemintArgs a = runs[0];
  const StepState ss = *a.st;
  const int N = a.N;
  double* accF = smem;
  double* accR = accF + N * MB;
  __shared__ double red[MEMINT_FINISH_THREADS / 64][MB];
  __shared__ double diagF[2][MB];
  const int mb = blockIdx.x;
  const double* in = part + (size_t)mb * nsplit * 2 * N * MB;
  for (int i = threadIdx.x; i < 2 * N * MB; i += blockDim.x) {
    double v = 0.0;
    for (int k = 0; k < nsplit; ++k) v += in[(size_t)k * 2 * N * MB + i];
    smem[i] = v;
  }
  __syncthreads();
  memint_leapfrog(a, ss, mb, accF, accR, red, diagF);
}

// the tile shapes the autotuner chooses from, as the whole-block kernel and
// the split form's partial kernel (the item table in MemintArgs follows the
// context's shape); entry 0 is the default
#define KB_MEMINT_SHAPE(R, B, T) \
  { R, B, T, (const void*)memint_step<R, B, T>, (const void*)memint_partial<R, B, T> }
const MemintVariant kMemintVariants[NMEMINT_VARIANTS] = {
    KB_MEMINT_SHAPE(64, 2, 256), KB_MEMINT_SHAPE(32, 4, 256),  KB_MEMINT_SHAPE(128, 4, 256),
    KB_MEMINT_SHAPE(64, 4, 256), KB_MEMINT_SHAPE(128, 2, 256), KB_MEMINT_SHAPE(32, 2, 256),
    KB_MEMINT_SHAPE(64, 4, 512), KB_MEMINT_SHAPE(128, 2, 512), KB_MEMINT_SHAPE(32, 4, 512),
    KB_MEMINT_SHAPE(128, 1, 512), KB_MEMINT_SHAPE(64, 1, 512),  KB_MEMINT_SHAPE(128, 1, 256),
    KB_MEMINT_SHAPE(64, 1, 256),
};
#undef KB_MEMINT_SHAPE

// ---------------------------------------------------------------------------
// Measurement (createMeasurement / getE_kin / getE_pot / getE_mem)
// partial sums per workgroup: [0] sum ddF/2, [1] sum F, [2] sum kL2 F,
// [3] E_mem boundary integrand, [4] E_mem interior integrand, [5] n_unstable
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void measure_kernel(MeasArgs a) {
  __shared__ double red[6][256];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double v[6] = {0, 0, 0, 0, 0, 0};
  if (p < a.V) {
    const int mb = p / MB, lane = p % MB;
    const double dt = a.dt;
    const double Ftt = a.H[hoff(mb, pair_off(a.s_t, a.s_t, a.N), lane, a.NP)];
    const double Fmt = a.H[hoff(mb, pair_off(a.s_tm1, a.s_t, a.N), lane, a.NP)];
    const double* emm = a.H + hoff(mb, pair_off(a.s_tm1, a.s_tm1, a.N), lane, a.NP);
    const double Fmm = emm[0];
    const double ddF = (Ftt - 2.0 * Fmt + Fmm) / (dt * dt);
    const double om = (ddF / Fmm) > 0 ? sqrt(ddF / Fmm) : 0.0;
    double c;
    int bad = 0;
    if (0.25 * (dt * dt) * (om * om) < 1.0) {
      c = sqrt(1.0 - 0.25 * (dt * dt) * (om * om));
    } else {
      c = 1.0;
      bad = 1;
    }
    const double nn = (ddF * Fmm > 0) ? (c * sqrt(ddF * Fmm)) - 0.5 : 0.0;
    if (a.outF) {
      a.outF[p] = Fmm;
      a.outddF[p] = ddF;
      a.outr[p] = emm[1];
      a.outn[p] = nn;
      a.outom[p] = om;
      a.outc[p] = c;
    }
    const double g = a.wgt ? a.wgt[p] : 1.0;
    v[0] = g * 0.5 * ddF;
    v[1] = g * Fmm;
    v[2] = g * a.kL2[p] * Fmm;
    v[5] = g * bad;
    if (a.t != 2 && a.integral) {
      // E_mem integrand over t' in [tmin, t-2] against F(t',t-1), rho(t',t-1)
      for (int tp = a.tmin; tp <= a.t - 2; ++tp) {
        const int x = (tp - 1) % a.N;
        const double* e = a.H + hoff(mb, pair_off(x, a.s_tm1, a.N), lane, a.NP);
        const double rr = e[1] * (x >= a.s_tm1 ? 1.0 : -1.0);
        const double* S = a.S + soff(mb, x, lane, a.N);
        const double term = g * (S[1] * e[0] - S[0] * rr);
        if (tp == a.tmin)
          v[3] += term;
        else
          v[4] += term;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 6) a.partial[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// site-1 history (Measurement.jl:90-96); q = 0 .. t-tmin-1
__global__ void history_kernel(const double* __restrict__ H, const double* __restrict__ S, long NP, int N,
                               int tmin, int t, double* __restrict__ out) {
  const int L = t - tmin;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L) return;
  // SigF / Sigr for tp = t-1 down to tmin
  const int tpS = t - 1 - q;
  const int xS = (tpS - 1) % N;
  out[q] = S[soff(0, xS, 0, N)];
  out[L + q] = S[soff(0, xS, 0, N) + 1];
  // F(tmin,tp), rho(tmin,tp) for tp = tmin + q
  const int xm = (tmin - 1) % N;
  const int x = (tmin + q - 1) % N;
  const double* e = H + hoff(0, pair_off(xm, x, N), 0, NP);
  out[2 * L + q] = e[0];
  out[3 * L + q] = e[1] * (xm >= x ? 1.0 : -1.0);
}

// ---------------------------------------------------------------------------
// host-API helpers
// ---------------------------------------------------------------------------
// out[p] = comp(pair)[p] * sign for p < V
__global__ void gather_pair(const double* __restrict__ H, long NP, int V, long pair, int comp, double sign,
                            double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < V) out[p] = H[hoff(p / MB, pair, p % MB, NP) + comp] * sign;
}

__global__ void scatter_pair(double* __restrict__ H, long NP, int V, long pair, int comp, double sign,
                             const double* __restrict__ in) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < V) H[hoff(p / MB, pair, p % MB, NP) + comp] = in[p] * sign;
}

// Streaming-read probe (benchmark support): each thread reads U 16-byte
// vectors per iteration, non-temporally, and the block writes one sum.
__global__ __launch_bounds__(256) void read_probe_kernel(const double* __restrict__ x, size_t n2,
                                                          double* __restrict__ out) {
  typedef double rp_d2 __attribute__((ext_vector_type(2)));
  constexpr int U = 8;
  const rp_d2* v = reinterpret_cast<const rp_d2*>(x);
  double acc = 0.0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n2; i += U * stride) {
    rp_d2 a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = __builtin_nontemporal_load(v + i + u * stride);
#pragma unroll
    for (int u = 0; u < U; ++u) acc += a[u].x + a[u].y;
  }
  for (; i < n2; i += stride) acc += v[i].x + v[i].y;
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double sum = 0.0;
    for (int k = 0; k < (int)blockDim.x; ++k) sum += red[k];
    out[blockIdx.x] = sum;
  }
}

__global__ void gather_tri(const double* __restrict__ P, long PNP, int V, long pp, double sign,
                           double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < V) out[p] = P[poff(p / MB, pp, p % MB, PNP)] * sign;
}

__global__ void gather_sigma(const double* __restrict__ S, int N, int V, int slot, int comp,
                             double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < V) out[p] = S[soff(p / MB, slot, p % MB, N) + comp];
}

// initSimData! rows (Initialization.jl:8-29): F(1,1)=F(2,1)=(n+1/2)/w,
// F(2,2)=w(n+1/2)dt^2+(n+1/2)/w, rho(2,1)=dt, diagonal rho = 0
__global__ void init_rows_kernel(double* __restrict__ H, long NP, int V, int N, const double* __restrict__ F11,
                                 const double* __restrict__ F21, const double* __restrict__ F22, double dt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= V) return;
  const int mb = p / MB, lane = p % MB;
  const int s1 = 0, s2 = 1 % N;
  double* e11 = H + hoff(mb, pair_off(s1, s1, N), lane, NP);
  double* e21 = H + hoff(mb, pair_off(s2, s1, N), lane, NP);
  double* e22 = H + hoff(mb, pair_off(s2, s2, N), lane, NP);
  e11[0] = F11[p];
  e11[1] = 0.0;
  e21[0] = F21[p];
  e21[1] = dt * (s2 >= s1 ? 1.0 : -1.0);  // r[2,1] = dt * thesign(2,1)
  e22[0] = F22[p];
  e22[1] = 0.0;
}

__global__ void init_particles_kernel(const double* __restrict__ kL2, const double* __restrict__ n, double m,
                                      double dt, int V, double* __restrict__ F11, double* __restrict__ F21,
                                      double* __restrict__ F22) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= V) return;
  const double omega = sqrt(kL2[p] + m * m);
  const double np = n[p];
  F11[p] = (1.0 / omega) * (np + 0.5);
  F21[p] = (1.0 / omega) * (np + 0.5);
  F22[p] = (omega * (np + 0.5) * (dt * dt) + (1.0 / omega) * (np + 0.5));
}

// full deterministic window of free propagators (benchmark set-up);
// blockIdx.y = slot column c, threads over (site, slot row r >= c)
__global__ void fill_free_kernel(double* __restrict__ H, long NP, int V, int N, int tmax,
                                 const double* __restrict__ kL2, const double* __restrict__ n, double mass,
                                 double dt, int nb) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (gid >= (long)nb * MB * N) return;
  const int r = (int)(gid % N);
  const int p = (int)(gid / N);
  if (r < c) return;
  const int mb = p / MB, lane = p % MB;
  double F = 0.0, rho = 0.0;
  if (p < V) {
    // times of the slots in the window [tmax-N+1, tmax]
    const int tmin = tmax - N + 1;
    const int smin = (tmin - 1) % N;
    const int th = tmin + (r - smin + N) % N;
    const int tl = tmin + (c - smin + N) % N;
    const double w = sqrt(kL2[p] + mass * mass);
    const double nn = n ? n[p] : 0.0;
    const double d = (double)(th - tl) * dt;
    F = (nn + 0.5) / w * cos(w * d);
    rho = sin(w * d) / w;  // rho(th, tl): the stored value for slots (r, c)
  }
  double* e = H + hoff(mb, pair_off(r, c, N), lane, NP);
  e[0] = F;
  e[1] = rho;
}

__global__ void finite_kernel(const double* __restrict__ H, size_t count, int* __restrict__ flag) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  int bad = 0;
  for (; i < count; i += stride) bad |= !isfinite(H[i]);
  if (bad) atomicOr(flag, 1);
}

}  // namespace qftkb
