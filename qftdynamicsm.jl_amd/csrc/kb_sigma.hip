// kb_sigma.hip -- self-energy stage of the KB step on whole lattice planes
// (TwoPIScalar_Evolution.jl:16-31), for Nx <= 64 and 1D lines up to 2048;
// past them the 1/N bubble and chain passes run as pencil passes
// (sigma_pencil, at the end of the kernels).
//
// For every live t' the sunset needs Fx = real(rfft(F(t-1,t'))) and
// rx = real(rfft(rho(t-1,t'))), the pointwise products rx (Fx^2 - rx^2/12)
// and Fx (Fx^2 - 3/4 rx^2), and their inverse transforms.  real(rfft(X))
// only sees the inversion-even part X_e(p) = (X(p) + X(-p))/2, whose
// transform is real, so F_e + i rho_e is transformed as one complex lattice
// and both transforms come out as its real and imaginary parts.
//
//   3D: plane_fwd  gathers the z-planes iz and -iz of F, rho, symmetrises,
//                  2D-transforms plane iz in LDS and writes it;
//       zpass      z transform, sunset products, inverse z transform;
//       plane_inv  inverse 2D transform, scale, write SigF/Sigr.
//   Every partial transform W of an even lattice is even under the point
//   reflection of all three indices, W(-x, -y, -z) = W(x, y, z), so the
//   workspace holds only the half space z in [0, nz/2] of each row: plane_fwd
//   runs the planes of that half; zpass runs the y rows y in [0, ny/2], reads
//   the other half of its z lines from the reflected row -y and writes both
//   rows' half spaces; plane_inv runs the planes of the half and writes
//   plane -z of the output as the reflection of plane z.
//   2D / 1D: plane_fwd does everything for its row in one workgroup.
//
// The same passes run the adjoint of the stage (SG_SUNSET_ADJ): the input is
// the adjoint self-energy pair (SigF_bar, Sigr_bar) per t', the pointwise
// step is the transpose of the sunset's Jacobian at the saved forward
// transforms (Fx, rx), and the output is added to the adjoint history row
// t-1.  The cosine transform C (real part of the DFT of an even lattice) is
// symmetric, so Sigma = pre C g(C F, C rho) has the adjoint
//   g_bar = pre C Sigma_bar,  F_bar += C (dg/dFx)^T g_bar,  rho_bar += C (dg/drx)^T g_bar.
//
// Two more modes serve the 1/N expansion (NinverseNLO):
//   BUBBLE  gathers F(t-1,t'), rho(t-1,t') like the sunset, keeps (Fx, rx)
//           in X and transforms Fx^2 - rx^2/4 and Fx rx back: the bubbles
//           F2kr2 = ivol^2 C[Fx^2 - rx^2/4] and Fr = ivol^2 C[Fx rx] of the
//           pair (t-1, t') (calc_Fr_F2kr2!, SimDataStruct.jl:726-750), stored
//           as row t-1 of two slot triangles;
//   CHAIN   gathers the resummed kernels (IF, Ir)(t') in momentum space,
//           transforms them to Irx = ivol C Ir, IFx = ivol C IF and forms
//           Sigr = pre C[Fx Irx + rx IFx], SigF = pre C[Fx IFx - rx Irx / 4]
//           with the saved (Fx, rx) (calcSelfEnergyattp!, Evolution.jl:683-698).
//
// All transforms are unnormalised Stockham FFTs in LDS; lines of a plane are
// padded to Nx+1 elements so that column access does not hit one bank.
// The FFT length is a template parameter for the common sizes (0 = runtime).
#include "kb_kernels.h"

namespace qftkb {

namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 twid(const double2* __restrict__ tw, int k, bool inv) {
  double2 w = tw[k];
  if (inv) w.y = -w.y;
  return w;
}

struct Fac {
  int nf;
  int f[24];
};

constexpr Fac factorize(int n) {
  Fac F{0, {}};
  int m = n;
  while (m % 4 == 0) { F.f[F.nf++] = 4; m /= 4; }
  while (m % 2 == 0) { F.f[F.nf++] = 2; m /= 2; }
  for (int p = 3; m > 1; p += 2)
    while (m % p == 0) { F.f[F.nf++] = p; m /= p; }
  return F;
}

// i z (sign +1) or -i z (sign -1)
__device__ __forceinline__ double2 imul(double2 z, double sign) { return make_double2(-sign * z.y, sign * z.x); }

template <int R>
__device__ __forceinline__ void small_dft(double2 (&v)[R], const double2* __restrict__ tw, int n, bool inv) {
  if constexpr (R == 2) {
    const double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 3) {
    // w = exp(-+2 pi i / 3): X1,2 = x0 - (x1 + x2)/2 -+ i (sqrt 3 / 2)(x1 - x2), sign flipped for the inverse
    constexpr double S3 = 0.86602540378443864676;  // sin(2 pi / 3)
    const double sg = inv ? 1.0 : -1.0;
    const double2 t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    const double2 m = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
    const double2 r = imul(make_double2(S3 * t2.x, S3 * t2.y), sg);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  } else if constexpr (R == 5) {
    // w = exp(-+2 pi i / 5) with c1 = cos(2 pi/5), c2 = cos(4 pi/5), s1 = sin(2 pi/5),
    // s2 = sin(4 pi/5), t1 = x1 + x4, t2 = x2 + x3, t3 = x1 - x4, t4 = x2 - x3:
    //   X1,4 = x0 + c1 t1 + c2 t2 -+ i (s1 t3 + s2 t4)
    //   X2,3 = x0 + c2 t1 + c1 t2 -+ i (s2 t3 - s1 t4)   (i terms flipped for the inverse)
    constexpr double C1 = 0.30901699437494742410, C2 = -0.80901699437494742410;
    constexpr double S1 = 0.95105651629515357212, S2 = 0.58778525229247312917;
    const double sg = inv ? 1.0 : -1.0;
    const double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const double2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const double2 a1 = make_double2(v[0].x + C1 * t1.x + C2 * t2.x, v[0].y + C1 * t1.y + C2 * t2.y);
    const double2 a2 = make_double2(v[0].x + C2 * t1.x + C1 * t2.x, v[0].y + C2 * t1.y + C1 * t2.y);
    const double2 b1 = imul(make_double2(S1 * t3.x + S2 * t4.x, S1 * t3.y + S2 * t4.y), sg);
    const double2 b2 = imul(make_double2(S2 * t3.x - S1 * t4.x, S2 * t3.y - S1 * t4.y), sg);
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
  } else if constexpr (R == 4) {
    const double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const double2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
    const double2 jb1 = inv ? make_double2(-b1.y, b1.x) : make_double2(b1.y, -b1.x);
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

// One Stockham stage on `nlines` lines; element (line, j) at line*ls + j*es.
// NL > 0: nlines == NL, known at compile time.
template <int R, int NL = 0>
__device__ __forceinline__ void stage(const double2* __restrict__ src, double2* __restrict__ dst, int n, int Ns,
                                      int nlines, int ls, int es, const double2* __restrict__ tw, bool inv) {
  const int m = n / R;
  const int step = n / (Ns * R);
  const int nl = NL > 0 ? NL : nlines;
  // neighbouring lanes take the same butterfly of neighbouring lines: their
  // stride-R writes (and reads) then fall on different LDS banks, and they
  // share the twiddle
  for (int idx = threadIdx.x; idx < nl * m; idx += blockDim.x) {
    const int j = idx / nl;
    const int line = idx - j * nl;
    const int kk = j % Ns;
    const double2* s = src + line * ls;
    double2* d = dst + line * ls;
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double2 x = s[(j + r * m) * es];
      // the first stage (Ns = 1) has unit twiddles
      v[r] = (r == 0 || Ns == 1) ? x : cmul(x, twid(tw, r * kk * step, inv));  // r*kk*step < n
    }
    small_dft<R>(v, tw, n, inv);
    const int base = (j / Ns) * Ns * R + kk;
#pragma unroll
    for (int q = 0; q < R; ++q) d[(base + q * Ns) * es] = v[q];
  }
}

// prime radix > 7: one output element per thread
__device__ void stage_direct(const double2* __restrict__ src, double2* __restrict__ dst, int n, int Ns, int R,
                             int nlines, int ls, int es, const double2* __restrict__ tw, bool inv) {
  const int m = n / R;
  const int step = n / (Ns * R);
  for (int idx = threadIdx.x; idx < nlines * n; idx += blockDim.x) {
    const int rem = idx / nlines;
    const int line = idx - rem * nlines;
    const int j = rem / R;
    const int q = rem - j * R;
    const int kk = j % Ns;
    const double2* s = src + line * ls;
    double2 acc = make_double2(0.0, 0.0);
    for (int r = 0; r < R; ++r) {
      double2 x = s[(j + r * m) * es];
      if (r) x = cmul(x, twid(tw, r * kk * step, inv));
      acc = cadd(acc, cmul(x, twid(tw, ((q * r) % R) * m, inv)));
    }
    dst[line * ls + ((j / Ns) * Ns * R + kk + q * Ns) * es] = acc;
  }
}

template <int R, int NL = 0>
__device__ __forceinline__ void stage_any(const double2* src, double2* dst, int n, int Ns, int nlines, int ls, int es,
                                          const double2* tw, bool inv) {
  if constexpr (R == 2 || R == 3 || R == 4 || R == 5 || R == 7)
    stage<R, NL>(src, dst, n, Ns, nlines, ls, es, tw, inv);
  else
    stage_direct(src, dst, n, Ns, R, nlines, ls, es, tw, inv);
}

// compile-time length: every stage's radix, span and stride are constants
template <int N, int S, int NS>
__device__ __forceinline__ double2* stages_ct(double2* a, double2* b, int nlines, int ls, int es,
                                              const double2* tw, bool inv) {
  constexpr Fac F = factorize(N);
  if constexpr (S >= F.nf) {
    return a;
  } else {
    constexpr int R = F.f[S];
    if (nlines == N)  // the 3D planes and pencils: as many lines as points
      stage_any<R, N>(a, b, N, NS, nlines, ls, es, tw, inv);
    else
      stage_any<R>(a, b, N, NS, nlines, ls, es, tw, inv);
    __syncthreads();
    return stages_ct<N, S + 1, NS * R>(b, a, nlines, ls, es, tw, inv);
  }
}

__device__ double2* stages_rt(double2* a, double2* b, const FFTPlan& pl, int nlines, int ls, int es,
                              const double2* tw, bool inv) {
  int Ns = 1;
  const int n = pl.n;
  for (int s = 0; s < pl.nf; ++s) {
    const int R = pl.fac[s];
    switch (R) {
      case 2: stage<2>(a, b, n, Ns, nlines, ls, es, tw, inv); break;
      case 3: stage<3>(a, b, n, Ns, nlines, ls, es, tw, inv); break;
      case 4: stage<4>(a, b, n, Ns, nlines, ls, es, tw, inv); break;
      case 5: stage<5>(a, b, n, Ns, nlines, ls, es, tw, inv); break;
      case 7: stage<7>(a, b, n, Ns, nlines, ls, es, tw, inv); break;
      default: stage_direct(a, b, n, Ns, R, nlines, ls, es, tw, inv); break;
    }
    __syncthreads();
    double2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// Transform lines held in a (ping-pong with b); returns the result buffer.
template <int NT>
__device__ __forceinline__ double2* fft_lines(double2* a, double2* b, const FFTPlan& pl, int nlines, int ls, int es,
                                              const double2* tw, bool inv) {
  if constexpr (NT > 0)
    return stages_ct<NT, 0, 1>(a, b, nlines, ls, es, tw, inv);
  else
    return stages_rt(a, b, pl, nlines, ls, es, tw, inv);
}

// (Fx (Fx^2 - qF rx^2), rx (Fx^2 - qr rx^2)): qF = 3/4, qr = 1/12 for the
// quantum sunset (Evolution.jl:25-31), 0 for the classical one (:107-120)
__device__ __forceinline__ double2 sunset(double2 z, int classical) {
  const double qF = classical ? 0.0 : 0.75, qr = classical ? 0.0 : 1.0 / 12.0;
  const double Fx = z.x, rx = z.y;
  const double FF = Fx * Fx, rr = rx * rx;
  return make_double2(Fx * (FF - qF * rr), rx * (FF - qr * rr));
}

// transpose of the sunset's Jacobian at (Fx, rx) applied to
// (gF, gr) = (preF, preR) * (C SigF_bar, C Sigr_bar)
__device__ __forceinline__ double2 sunset_adj(double2 cs, double2 x, double preF, double preR, int classical) {
  const double a1 = classical ? 0.0 : 0.75, a2 = classical ? 0.0 : -1.5, a3 = classical ? 0.0 : 0.25;
  const double gF = preF * cs.x, gr = preR * cs.y;
  const double Fx = x.x, rx = x.y;
  const double FF = Fx * Fx, rr = rx * rx;
  return make_double2(gF * (3.0 * FF - a1 * rr) + gr * (2.0 * rx * Fx), gF * (a2 * Fx * rx) + gr * (FF - a3 * rr));
}

// the pointwise step of each mode; the sunset and bubble save (Fx, rx) in X when it is set
template <int MODE>
__device__ __forceinline__ double2 pointwise(const SigmaArgs& a, double2 z, size_t xi) {
  if constexpr (MODE == SG_SUNSET_ADJ) {
    return sunset_adj(z, a.X[xi], a.preF, a.preR, a.classical);
  } else if constexpr (MODE == SG_BUBBLE) {
    if (a.X) a.X[xi] = z;
    return make_double2(z.x * z.x - 0.25 * (z.y * z.y), z.x * z.y);
  } else if constexpr (MODE == SG_CHAIN) {
    const double2 x = a.X[xi];  // (Fx, rx) of (t-1, t')
    const double IFx = a.ivol * z.x, Irx = a.ivol * z.y;
    return make_double2(x.x * IFx - 0.25 * (x.y * Irx), x.x * Irx + x.y * IFx);
  } else {
    if (a.X) a.X[xi] = z;
    return sunset(z, a.classical);
  }
}

// output of the stage for row q (time t'), site s: Sigma into the ring, the
// adjoint added to (F_bar, rho_bar)(t-1, t'), or the bubbles of (t-1, t')
template <int MODE>
__device__ __forceinline__ void emit(const SigmaArgs& a, int s, int slot, size_t rowoff, double2 z) {
  if constexpr (MODE == SG_PLAIN) {
    a.SigFout[rowoff + s] = a.preF * z.x;
    a.Sigrout[rowoff + s] = a.preR * z.y;
  } else if constexpr (MODE == SG_SUNSET_ADJ) {
    const int s_tm1 = a.st->s_tm1;
    double* o = a.Hbar + hoff(s / MB, pair_off(s_tm1, slot, a.N), s % MB, a.NP);
    o[0] += z.x;
    if (slot != s_tm1) o[1] += (s_tm1 >= slot ? 1.0 : -1.0) * z.y;  // rho(t-1,t-1) is not a variable
  } else {
    if (a.rmap) {
      s = a.rmap[s];  // orbit storage: the representative writes its orbit's value
      if (s < 0) return;
    }
    if constexpr (MODE == SG_BUBBLE) {
      const int s_tm1 = a.st->s_tm1;
      const size_t o = poff(s / MB, ppair_off(s_tm1, slot, a.N), s % MB, a.PNP);
      a.PF[o] = a.preF * z.x;
      a.PR[o] = (s_tm1 >= slot ? a.preR : -a.preR) * z.y;  // sign-folded like rho
    } else {
      double* o = a.S + soff(s / MB, slot, s % MB, a.N);
      o[0] = a.preF * z.x;
      o[1] = a.preR * z.y;
    }
  }
}

// rows of the stage: the live times t' = tmin + q, q < L, of the window, or
// the caller's lattices
template <int MODE>
__device__ __forceinline__ int stage_rows(const SigmaArgs& a) {
  if constexpr (MODE == SG_PLAIN)
    return a.rows;
  else
    return a.st->L;
}

// ring slot of row q (unused for caller lattices)
template <int MODE>
__device__ __forceinline__ int row_slot(const SigmaArgs& a, int q) {
  if constexpr (MODE == SG_PLAIN)
    return 0;
  else
    return (a.st->tmin + q - 1) % a.N;
}

}  // namespace

// ---------------------------------------------------------------------------
// blockIdx.x: plane iz (3D: 0 .. nz/2), blockIdx.y: row q (t' = tmin + q)
template <int NT, int MODE, int TPB>
__global__ __launch_bounds__(TPB) void sigma_plane_fwd(SigmaArgs a) {
  extern __shared__ double2 lds[];
  const int n = NT > 0 ? NT : a.n;
  const int ny = a.ny, nz = a.nz;
  const int ls = n + 1;  // padded line
  double2* A = lds;
  double2* B = lds + ny * ls;
  const int q = blockIdx.y;
  if (q >= stage_rows<MODE>(a)) return;
  const int iz = blockIdx.x;
  const int izn = (nz - iz) % nz;
  const int slot = row_slot<MODE>(a, q);
  const size_t rowoff = (size_t)q * a.V;
  const int nplane = n * ny;

  // physical F(t-1,t'), rho(t-1,t') (adjoint: SigF_bar, Sigr_bar at t') of
  // plane iz and the reflected plane -iz, inversion-symmetrised
  for (int e = threadIdx.x; e < nplane; e += blockDim.x) {
    const int iy = e / n, ix = e - iy * n;
    const int s1 = ix + n * (iy + ny * iz);
    const int s2 = (n - ix) % n + n * ((ny - iy) % ny + ny * izn);
    if constexpr (MODE == SG_PLAIN) {
      const double* F = a.Fin + rowoff;
      const double* R = a.Rin + rowoff;
      A[iy * ls + ix] = make_double2(0.5 * (F[s1] + F[s2]), 0.5 * (R[s1] + R[s2]));
    } else if constexpr (MODE == SG_SUNSET_ADJ || MODE == SG_CHAIN) {
      // a ring in self-energy layout: (SigF_bar, Sigr_bar) or (IF, Ir)
      const double* ring = MODE == SG_CHAIN ? a.Iring : a.Sbar;
      const int q1 = a.smap ? a.smap[s1] : s1, q2 = a.smap ? a.smap[s2] : s2;
      const double* h1 = ring + soff(q1 / MB, slot, q1 % MB, a.N);
      const double* h2 = ring + soff(q2 / MB, slot, q2 % MB, a.N);
      A[iy * ls + ix] = make_double2(0.5 * (h1[0] + h2[0]), 0.5 * (h1[1] + h2[1]));
    } else {
      const int s_tm1 = a.st->s_tm1;
      const long pr = pair_off(s_tm1, slot, a.N);
      const double sg = s_tm1 >= slot ? 1.0 : -1.0;
      const int q1 = a.smap ? a.smap[s1] : s1, q2 = a.smap ? a.smap[s2] : s2;
      const double* h1 = a.H + hoff(q1 / MB, pr, q1 % MB, a.NP);
      const double* h2 = a.H + hoff(q2 / MB, pr, q2 % MB, a.NP);
      A[iy * ls + ix] = make_double2(0.5 * (h1[0] + h2[0]), 0.5 * (h1[1] + h2[1]) * sg);
    }
  }
  __syncthreads();
  double2* res = fft_lines<NT>(A, B, a.plan, ny, ls, 1, a.tw, false);
  double2* oth = res == A ? B : A;
  if (ny > 1) {
    res = fft_lines<NT>(res, oth, a.plan, n, 1, ls, a.tw, false);
    oth = res == A ? B : A;
  }
  if (nz == 1) {
    // 1D / 2D: the whole lattice is in LDS -- sunset, inverse, write Sigma
    for (int e = threadIdx.x; e < nplane; e += blockDim.x) {
      const int iy = e / n, ix = e - iy * n;
      res[iy * ls + ix] = pointwise<MODE>(a, res[iy * ls + ix], rowoff + e);
    }
    __syncthreads();
    if (ny > 1) {
      res = fft_lines<NT>(res, oth, a.plan, n, 1, ls, a.tw, true);
      oth = res == A ? B : A;
    }
    res = fft_lines<NT>(res, oth, a.plan, ny, ls, 1, a.tw, true);
    for (int e = threadIdx.x; e < nplane; e += blockDim.x) {
      const int iy = e / n, ix = e - iy * n;
      emit<MODE>(a, ix + n * iy, slot, rowoff, res[iy * ls + ix]);
    }
    return;
  }
  const size_t wrow = (size_t)q * nplane * (nz / 2 + 1);  // half-space workspace row
  for (int e = threadIdx.x; e < nplane; e += blockDim.x) {
    const int iy = e / n, ix = e - iy * n;
    a.W[wrow + ix + n * (iy + ny * iz)] = res[iy * ls + ix];
  }
}

// blockIdx.x: y row iy in [0, ny/2], blockIdx.y: row q.  Lines along z for
// every x; the workgroup owns the workspace rows iy and -iy.
template <int NT, int MODE, int TPB>
__global__ __launch_bounds__(TPB) void sigma_zpass(SigmaArgs a) {
  extern __shared__ double2 lds[];
  const int n = NT > 0 ? NT : a.n;
  const int ny = a.ny, nz = a.nz;
  const int ls = nz + 1;
  const int nzh = nz / 2 + 1;
  double2* A = lds;
  double2* B = lds + n * ls;
  const int q = blockIdx.y;
  if (q >= stage_rows<MODE>(a)) return;
  const int iy = blockIdx.x;
  const int iyn = (ny - iy) % ny;
  const size_t rowoff = (size_t)q * a.V;               // site-indexed arrays (X)
  const size_t wrow = (size_t)q * n * ny * nzh;        // half-space workspace row
  for (int e = threadIdx.x; e < n * nz; e += blockDim.x) {
    const int iz = e / n, ix = e - iz * n;
    // z >= nzh: W(x, y, z) = W(-x, -y, -z), from the half space of row -y
    A[ix * ls + iz] = iz < nzh ? a.W[wrow + ix + n * (iy + ny * iz)]
                               : a.W[wrow + (n - ix) % n + n * (iyn + ny * (nz - iz))];
  }
  __syncthreads();
  double2* res = fft_lines<NT>(A, B, a.plan, n, ls, 1, a.tw, false);
  double2* oth = res == A ? B : A;
  for (int e = threadIdx.x; e < n * nz; e += blockDim.x) {
    const int ix = e / nz, iz = e - ix * nz;
    res[ix * ls + iz] = pointwise<MODE>(a, res[ix * ls + iz], rowoff + ix + n * (iy + ny * iz));
  }
  __syncthreads();
  res = fft_lines<NT>(res, oth, a.plan, n, ls, 1, a.tw, true);
  for (int e = threadIdx.x; e < n * nzh; e += blockDim.x) {
    const int iz = e / n, ix = e - iz * n;
    a.W[wrow + ix + n * (iy + ny * iz)] = res[ix * ls + iz];
    if (iyn != iy) a.W[wrow + ix + n * (iyn + ny * iz)] = res[((n - ix) % n) * ls + (nz - iz) % nz];
  }
}

// blockIdx.x: plane iz in [0, nz/2], blockIdx.y: row q.  Inverse 2D
// transform + scatter of planes iz and -iz.
template <int NT, int MODE, int TPB>
__global__ __launch_bounds__(TPB) void sigma_plane_inv(SigmaArgs a) {
  extern __shared__ double2 lds[];
  const int n = NT > 0 ? NT : a.n;
  const int ny = a.ny;
  const int ls = n + 1;
  double2* A = lds;
  double2* B = lds + ny * ls;
  const int q = blockIdx.y;
  if (q >= stage_rows<MODE>(a)) return;
  const int nz = a.nz;
  const int iz = blockIdx.x;
  const int izn = (nz - iz) % nz;
  const int slot = row_slot<MODE>(a, q);
  const size_t rowoff = (size_t)q * a.V;
  const int nplane = n * ny;
  const size_t wrow = (size_t)q * nplane * (nz / 2 + 1);  // half-space workspace row
  for (int e = threadIdx.x; e < nplane; e += blockDim.x) {
    const int iy = e / n, ix = e - iy * n;
    A[iy * ls + ix] = a.W[wrow + ix + n * (iy + ny * iz)];
  }
  __syncthreads();
  double2* res = fft_lines<NT>(A, B, a.plan, n, 1, ls, a.tw, true);
  double2* oth = res == A ? B : A;
  res = fft_lines<NT>(res, oth, a.plan, ny, ls, 1, a.tw, true);
  for (int e = threadIdx.x; e < nplane; e += blockDim.x) {
    const int iy = e / n, ix = e - iy * n;
    emit<MODE>(a, ix + n * (iy + ny * iz), slot, rowoff, res[iy * ls + ix]);
    // plane -iz is the point reflection of plane iz
    if (izn != iz) emit<MODE>(a, ix + n * (iy + ny * izn), slot, rowoff, res[((ny - iy) % ny) * ls + (n - ix) % n]);
  }
}

// ---------------------------------------------------------------------------
// The 1/N bubble and chain passes, and the adjoint's sunset with its saved
// transforms, on 2D and 3D lattices past the plane kernels (Nx > 64), as
// pencil passes: a workgroup transforms P pencils of n
// points along a.axis of row q (blockIdx.y), and whole rows of W hold the
// partial transforms between passes.  a.flags:
//   PEN_GATHER  first pass: the mode's input, inversion-symmetrised
//   PEN_POINT   forward transform, pointwise step, inverse transform (the
//               last axis)
//   PEN_INV     the transform is the inverse one
//   PEN_EMIT    last pass: the mode's output instead of W
// A pass with neither PEN_POINT nor PEN_INV is a forward transform.
// ---------------------------------------------------------------------------
enum { PEN_GATHER = 1, PEN_POINT = 2, PEN_INV = 4, PEN_EMIT = 8 };

namespace {

// site of point j of pencil `pencil` along a.axis (pencils enumerate the two
// other coordinates, the lower one fastest)
__device__ __forceinline__ int pencil_site(const SigmaArgs& a, int pencil, int j) {
  if (a.axis == 0) {
    const int iy = pencil % a.ny, iz = pencil / a.ny;
    return j + a.nx * (iy + a.ny * iz);
  }
  if (a.axis == 1) {
    const int ix = pencil % a.nx, iz = pencil / a.nx;
    return ix + a.nx * (j + a.ny * iz);
  }
  const int ix = pencil % a.nx, iy = pencil / a.nx;
  return ix + a.nx * (iy + a.ny * j);
}

// the point reflection x -> -x of a site
__device__ __forceinline__ int reflected_site(const SigmaArgs& a, int s) {
  const int ix = s % a.nx, iy = (s / a.nx) % a.ny, iz = s / (a.nx * a.ny);
  return (a.nx - ix) % a.nx + a.nx * ((a.ny - iy) % a.ny + a.ny * ((a.nz - iz) % a.nz));
}

// the pass input of full storage, symmetrised over sites s1 and s2 = -s1:
// (IF, Ir)(t') (chain) or (SigF_bar, Sigr_bar)(t') (adjoint) from a ring, or
// physical F(t-1,t'), rho(t-1,t') from the history (sunset, bubble)
template <int MODE>
__device__ __forceinline__ double2 pencil_gather(const SigmaArgs& a, int s1, int s2, int slot) {
  if constexpr (MODE == SG_CHAIN || MODE == SG_SUNSET_ADJ) {
    const double* ring = MODE == SG_CHAIN ? a.Iring : a.Sbar;
    const double* h1 = ring + soff(s1 / MB, slot, s1 % MB, a.N);
    const double* h2 = ring + soff(s2 / MB, slot, s2 % MB, a.N);
    return make_double2(0.5 * (h1[0] + h2[0]), 0.5 * (h1[1] + h2[1]));
  } else {
    const int s_tm1 = a.st->s_tm1;
    const long pr = pair_off(s_tm1, slot, a.N);
    const double sg = s_tm1 >= slot ? 1.0 : -1.0;
    const double* h1 = a.H + hoff(s1 / MB, pr, s1 % MB, a.NP);
    const double* h2 = a.H + hoff(s2 / MB, pr, s2 % MB, a.NP);
    return make_double2(0.5 * (h1[0] + h2[0]), 0.5 * (h1[1] + h2[1]) * sg);
  }
}

}  // namespace

template <int MODE>
__global__ __launch_bounds__(256) void sigma_pencil(SigmaArgs a) {
  extern __shared__ double2 lds[];
  const int n = a.n, P = a.P;
  double2* A = lds;
  double2* B = lds + (size_t)P * n;
  const int q = blockIdx.y;
  if (q >= stage_rows<MODE>(a)) return;
  const int slot = row_slot<MODE>(a, q);
  const size_t rowoff = (size_t)q * a.V;
  const int pen0 = blockIdx.x * P;
  // along x a pencil is contiguous: neighbouring threads take neighbouring
  // points; along y and z neighbouring pencils are adjacent in memory
  const bool contiguous = a.axis == 0;
  for (int e = threadIdx.x; e < P * n; e += blockDim.x) {
    const int k = contiguous ? e / n : e % P;
    const int j = contiguous ? e - k * n : e / P;
    const int pencil = pen0 + k;
    double2 z = make_double2(0.0, 0.0);
    if (pencil < a.npencils) {
      const int s = pencil_site(a, pencil, j);
      z = (a.flags & PEN_GATHER) ? pencil_gather<MODE>(a, s, reflected_site(a, s), slot) : a.W[rowoff + s];
    }
    A[k * n + j] = z;
  }
  __syncthreads();
  double2* res = fft_lines<0>(A, B, a.plan, P, n, 1, a.tw, (a.flags & PEN_INV) != 0);
  if (a.flags & PEN_POINT) {
    double2* oth = res == A ? B : A;
    for (int e = threadIdx.x; e < P * n; e += blockDim.x) {
      const int k = e / n, j = e - k * n;
      if (pen0 + k < a.npencils) res[e] = pointwise<MODE>(a, res[e], rowoff + pencil_site(a, pen0 + k, j));
    }
    __syncthreads();
    res = fft_lines<0>(res, oth, a.plan, P, n, 1, a.tw, true);
  }
  for (int e = threadIdx.x; e < P * n; e += blockDim.x) {
    const int k = contiguous ? e / n : e % P;
    const int j = contiguous ? e - k * n : e / P;
    const int pencil = pen0 + k;
    if (pencil >= a.npencils) continue;
    const int s = pencil_site(a, pencil, j);
    if (a.flags & PEN_EMIT)
      emit<MODE>(a, s, slot, rowoff, res[k * n + j]);
    else
      a.W[rowoff + s] = res[k * n + j];
  }
}

// ---------------------------------------------------------------------------
// host-side launch of the whole stage
// ---------------------------------------------------------------------------
template <int NT, int MODE, int TPB>
static hipError_t launch_planes_t(const SigmaArgs& a, int L, hipStream_t st) {  // L: grid rows
  const int n = a.n;
  const size_t lds_plane = sizeof(double2) * 2 * (size_t)a.ny * (n + 1);
  if (a.nz == 1) {
    hipLaunchKernelGGL((sigma_plane_fwd<NT, MODE, TPB>), dim3(1, L), dim3(TPB), lds_plane, st, a);
    return hipGetLastError();
  }
  const size_t lds_z = sizeof(double2) * 2 * (size_t)n * (a.nz + 1);
  hipLaunchKernelGGL((sigma_plane_fwd<NT, MODE, TPB>), dim3(a.nz / 2 + 1, L), dim3(TPB), lds_plane, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((sigma_zpass<NT, MODE, TPB>), dim3(a.ny / 2 + 1, L), dim3(TPB), lds_z, st, a);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((sigma_plane_inv<NT, MODE, TPB>), dim3(a.nz / 2 + 1, L), dim3(TPB), lds_plane, st, a);
  return hipGetLastError();
}

// planes of up to 64 x 64 points; a 1D lattice is one line, whose two LDS
// buffers (2 (Nx + 1) complex values, 64 KiB at Nx = 2048) fit the dynamic
// LDS that sigma_planes_prepare allows for a 64 x 64 plane
bool sigma_planes_supported(int Nx, int sdim) { return Nx >= 1 && Nx <= (sdim == 1 ? 2048 : 64); }

template <int NT, int MODE, int TPB>
static hipError_t prepare_one(size_t lds) {
  hipError_t e = hipFuncSetAttribute((const void*)sigma_plane_fwd<NT, MODE, TPB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void*)sigma_zpass<NT, MODE, TPB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void*)sigma_plane_inv<NT, MODE, TPB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

// threads per workgroup of the plane passes that the autotuner does not
// choose (the adjoint, 1/N bubble and chain, and context-free passes): the
// sunset passes' measured best at 32^3-48^3 (170 against 200 us per stage at
// 32^3, N_mem = 400, for 512 against 256 threads)
constexpr int PLANE_THREADS = 512;

template <int NT>
static hipError_t prepare_t(size_t lds) {
  hipError_t e = prepare_one<NT, SG_SUNSET, 256>(lds);
  if (e == hipSuccess) e = prepare_one<NT, SG_SUNSET, 128>(lds);
  if (e == hipSuccess) e = prepare_one<NT, SG_SUNSET, 512>(lds);
  if (e == hipSuccess) e = prepare_one<NT, SG_SUNSET, 1024>(lds);
  if (e == hipSuccess) e = prepare_one<NT, SG_BUBBLE, PLANE_THREADS>(lds);
  if (e == hipSuccess) e = prepare_one<NT, SG_CHAIN, PLANE_THREADS>(lds);
  if (e == hipSuccess) e = prepare_one<NT, SG_PLAIN, PLANE_THREADS>(lds);
  if (e != hipSuccess) return e;
  return prepare_one<NT, SG_SUNSET_ADJ, PLANE_THREADS>(lds);
}

// allow the dynamic LDS of the plane kernels for lattice size n (the
// attribute is per function and shared by all contexts: use the bound for
// the largest supported plane, Nx = 64)
hipError_t sigma_planes_prepare(int n) {
  const size_t lds = sizeof(double2) * 2 * (size_t)64 * (64 + 1);
  switch (n) {
    case 8: return prepare_t<8>(lds);
    case 16: return prepare_t<16>(lds);
    case 24: return prepare_t<24>(lds);
    case 32: return prepare_t<32>(lds);
    case 40: return prepare_t<40>(lds);
    case 48: return prepare_t<48>(lds);
    case 64: return prepare_t<64>(lds);
    default: return prepare_t<0>(lds);
  }
}

template <int MODE, int TPB>
static hipError_t launch_planes_n(const SigmaArgs& a, int L, hipStream_t st) {  // L: grid rows
  switch (a.n) {
    case 8: return launch_planes_t<8, MODE, TPB>(a, L, st);
    case 16: return launch_planes_t<16, MODE, TPB>(a, L, st);
    case 24: return launch_planes_t<24, MODE, TPB>(a, L, st);
    case 32: return launch_planes_t<32, MODE, TPB>(a, L, st);
    case 40: return launch_planes_t<40, MODE, TPB>(a, L, st);
    case 48: return launch_planes_t<48, MODE, TPB>(a, L, st);
    case 64: return launch_planes_t<64, MODE, TPB>(a, L, st);
    default: return launch_planes_t<0, MODE, TPB>(a, L, st);
  }
}

// the plane kernels' block sizes the autotuner chooses from; entry 0 is the default
const int kFFTThreads[NFFT_VARIANTS] = {256, 128, 512, 1024};

hipError_t launch_sigma_planes(const SigmaArgs& a, int L, hipStream_t st, int variant) {
  switch (variant) {
    case 1: return launch_planes_n<SG_SUNSET, 128>(a, L, st);
    case 2: return launch_planes_n<SG_SUNSET, 512>(a, L, st);
    case 3: return launch_planes_n<SG_SUNSET, 1024>(a, L, st);
    default: return launch_planes_n<SG_SUNSET, 256>(a, L, st);
  }
}

hipError_t launch_sigma_planes_adjoint(const SigmaArgs& a, int L, hipStream_t st) {
  return launch_planes_n<SG_SUNSET_ADJ, PLANE_THREADS>(a, L, st);
}

hipError_t launch_sigma_planes_plain(const SigmaArgs& a, hipStream_t st) {
  return launch_planes_n<SG_PLAIN, PLANE_THREADS>(a, a.rows, st);
}

hipError_t launch_bubble_planes(const SigmaArgs& a, int L, hipStream_t st) {
  return launch_planes_n<SG_BUBBLE, PLANE_THREADS>(a, L, st);
}

hipError_t launch_chain_planes(const SigmaArgs& a, int L, hipStream_t st) {
  return launch_planes_n<SG_CHAIN, PLANE_THREADS>(a, L, st);
}

// the pencil passes of one mode over 2D or 3D rows: forward along x (and y
// in 3D), forward + pointwise + inverse along the last axis, inverse back
template <int MODE>
static hipError_t launch_pencils(SigmaArgs a, int L, hipStream_t st) {
  if (a.ny == 1 || a.P < 1 || (size_t)a.P * a.n * 2 * sizeof(double2) > 65536) return hipErrorInvalidValue;
  const size_t lds = sizeof(double2) * 2 * (size_t)a.P * a.n;
  const dim3 grid((a.npencils + a.P - 1) / a.P, L);
  const bool three = a.nz > 1;
  const int passes[5][2] = {{0, PEN_GATHER}, {1, 0}, {three ? 2 : 1, PEN_POINT}, {1, PEN_INV}, {0, PEN_INV | PEN_EMIT}};
  for (int i = 0; i < 5; ++i) {
    if (!three && (i == 1 || i == 3)) continue;
    a.axis = passes[i][0];
    a.flags = passes[i][1];
    hipLaunchKernelGGL((sigma_pencil<MODE>), grid, dim3(256), lds, st, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_bubble_pencils(const SigmaArgs& a, int L, hipStream_t st) { return launch_pencils<SG_BUBBLE>(a, L, st); }

hipError_t launch_chain_pencils(const SigmaArgs& a, int L, hipStream_t st) { return launch_pencils<SG_CHAIN>(a, L, st); }

hipError_t launch_sunset_pencils(const SigmaArgs& a, int L, hipStream_t st) { return launch_pencils<SG_SUNSET>(a, L, st); }

hipError_t launch_sunset_adjoint_pencils(const SigmaArgs& a, int L, hipStream_t st) {
  return launch_pencils<SG_SUNSET_ADJ>(a, L, st);
}

}  // namespace qftkb
