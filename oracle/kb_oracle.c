/*
 * kb_oracle.c -- CPU restatement of the reference 2PI Kadanoff-Baym step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
 * CPU baseline reported beside it.  See kb_oracle.h for conventions.
 *
 * Every function cites the reference file:line it restates (paths are
 * relative to the reference repository root).  The storage and loop
 * structure deliberately follow the reference CPUfull back-end, including
 * its double read of the two-time history (getRHS_r! and getRHS_F! each
 * walk their own column/row of F and r), so that the baseline timing is
 * that of the reference algorithm.
 *
 * FFT: the reference calls FFTW rfft/brfft (TwoPIScalar_SimDataStruct.jl:
 * 381-382).  FFTW is not available here, so this file carries its own
 * mixed-radix Stockham FFT.  real(rfft(X)) equals the real part of the full
 * complex DFT restricted to the half grid, and brfft of the real, even
 * product equals the real part of the full inverse DFT; the oracle
 * therefore works on the full grid with complex transforms, which is the
 * same mathematics (parity of the transform itself is pinned against numpy
 * and against a direct double convolution in the tests).
 */
#include "kb_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  double re, im;
} cplx;

/* zeroed allocation for the large history arrays: anonymous mapping with
 * transparent huge pages requested (first-touch faults dominate otherwise) */
static void *big_alloc(size_t bytes) {
  if (bytes == 0) bytes = 8;
  void *p = mmap(NULL, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
  if (p == MAP_FAILED) return NULL;
  madvise(p, bytes, MADV_HUGEPAGE);
  return p;
}

/* ------------------------------------------------------------------ */
/* 1D mixed-radix Stockham FFT                                         */
/* ------------------------------------------------------------------ */
typedef struct {
  int n;
  int nf;
  int fac[40];
  cplx *w; /* w[k] = exp(-2 pi i k / n) */
} plan1d;

static void plan1d_init(plan1d *p, int n) {
  p->n = n;
  p->nf = 0;
  int m = n;
  while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
  while (m % 3 == 0) { p->fac[p->nf++] = 3; m /= 3; }
  while (m % 5 == 0) { p->fac[p->nf++] = 5; m /= 5; }
  for (int f = 7; m > 1; f += 2) {
    while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
  }
  p->w = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
  for (int k = 0; k < n; ++k) {
    /* exact in long double before rounding */
    long double a = -2.0L * 3.14159265358979323846264338327950288L *
                    (long double)k / (long double)n;
    p->w[k].re = (double)cosl(a);
    p->w[k].im = (double)sinl(a);
  }
}

static void plan1d_free(plan1d *p) { free(p->w); }

/* in-place transform of x (length n) using tmp (length n).  sign=-1 forward
 * exp(-i...), sign=+1 backward exp(+i...), both unnormalised. */
static void fft1d(const plan1d *p, cplx *x, cplx *tmp, int sign) {
  const int n = p->n;
  if (n <= 1) return;
  cplx *src = x, *dst = tmp;
  int Ns = 1;
  /* one butterfly's inputs and outputs; a prime factor above 64 (Nx = 67,
   * 71, ... on the pencil path) takes heap buffers */
  int maxR = 1;
  for (int s = 0; s < p->nf; ++s) maxR = p->fac[s] > maxR ? p->fac[s] : maxR;
  cplx vbuf[64], obuf[64];
  cplx *v = vbuf, *o = obuf;
  if (maxR > 64) {
    v = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)maxR);
    o = v + maxR;
  }
  for (int s = 0; s < p->nf; ++s) {
    const int R = p->fac[s];
    const int m = n / R;
    for (int j = 0; j < m; ++j) {
      const int k = j % Ns;
      for (int r = 0; r < R; ++r) {
        cplx a = src[j + r * m];
        /* twiddle exp(sign 2 pi i r k / (Ns R)) = w[r k m/Ns ... ] */
        long tw = ((long)r * k * (n / (Ns * R))) % n;
        cplx w = p->w[tw];
        if (sign > 0) w.im = -w.im;
        v[r].re = a.re * w.re - a.im * w.im;
        v[r].im = a.re * w.im + a.im * w.re;
      }
      for (int q = 0; q < R; ++q) {
        double sr = 0.0, si = 0.0;
        for (int r = 0; r < R; ++r) {
          long idx = ((long)q * r % R) * m; /* exp(-2 pi i q r / R) */
          cplx w = p->w[idx];
          if (sign > 0) w.im = -w.im;
          sr += v[r].re * w.re - v[r].im * w.im;
          si += v[r].re * w.im + v[r].im * w.re;
        }
        o[q].re = sr;
        o[q].im = si;
      }
      const int d = (j / Ns) * Ns * R + k;
      for (int q = 0; q < R; ++q) dst[d + q * Ns] = o[q];
    }
    cplx *t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  if (src != x) memcpy(x, src, sizeof(cplx) * (size_t)n);
  if (v != vbuf) free(v);
}

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */
struct or_ctx {
  int Nx, sdim, V, N;
  double dt, lambda, Mass, Mass2;
  int ON, pexp;
  int tmin, tmax;
  double *F, *r;       /* N(N+1)/2 lattices each, slot triangle */
  size_t bigbytes;
  double *SigF, *Sigr; /* N lattices each */
  /* NinverseNLO: Irx, IFx per slot (CyclVector of rfftlattice,
   * SimDataStruct.jl:70-71), kept as the complex full-grid extension */
  cplx *Irx, *IFx;
  long long ntransforms; /* full-lattice transforms of the 1/N stage so far */
  double *kL2, *omega2;
  double hartreemass2;
  int nthreads;
  plan1d plan;
  char err[256];
};

static int fail(or_ctx *c, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(c->err, sizeof c->err, fmt, ap);
  va_end(ap);
  return -1;
}

const char *or_last_error(or_ctx *c) { return c ? c->err : "null context"; }

static inline int mod1(int i, int N) { return ((i - 1) % N + N) % N + 1; }

/* thesign, DefSimDataElements.jl:224-230 */
static inline double thesign(int i, int j, int N) {
  return mod1(i, N) >= mod1(j, N) ? 1.0 : -1.0;
}

static inline int inwin(const or_ctx *c, int i) {
  return c->tmin <= i && i <= c->tmax;
}

/* SymMatrix getindex, DefSimDataElements.jl:185-195 */
static double *symptr(or_ctx *c, double *base, int i, int j) {
  int mi = mod1(i, c->N), mj = mod1(j, c->N);
  size_t idx;
  if (mi >= mj)
    idx = (size_t)mi * (mi - 1) / 2 + (mj - 1);
  else
    idx = (size_t)mj * (mj - 1) / 2 + (mi - 1);
  return base + idx * (size_t)c->V;
}

/* CyclVector getindex, DefSimDataElements.jl:129-132 */
static double *cycptr(or_ctx *c, double *base, int i) {
  return base + (size_t)(mod1(i, c->N) - 1) * (size_t)c->V;
}

static void compute_kL2(or_ctx *c) {
  const int Nx = c->Nx;
  for (int idx = 0; idx < c->V; ++idx) {
    int rem = idx;
    double k2 = 0.0;
    for (int d = 0; d < c->sdim; ++d) {
      int nd = rem % Nx;
      rem /= Nx;
      double s = sin(M_PI * (double)nd / (double)Nx);
      k2 += 4.0 * s * s;
    }
    c->kL2[idx] = k2;
  }
}

/* getTwoPISimData, SimDataStruct.jl:673-679 */
static void reset_omega2(or_ctx *c) {
  for (int i = 0; i < c->V; ++i) {
    c->omega2[i] = c->kL2[i] + c->Mass * c->Mass;
    c->omega2[i] *= c->dt * c->dt;
  }
}

or_ctx *or_create(int Nx, int sdim, int Nmem, double dt, double lambda,
                  int ONgroup, double Mass, double Mass2, int pexp,
                  int nthreads) {
  if (Nx < 1 || sdim < 1 || sdim > 3 || Nmem < 3 || ONgroup < 1 || pexp < 0 || pexp > 4) return NULL;
  or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
  c->Nx = Nx;
  c->sdim = sdim;
  c->V = 1;
  for (int d = 0; d < sdim; ++d) c->V *= Nx;
  c->N = Nmem;
  c->dt = dt;
  c->lambda = lambda;
  c->ON = ONgroup;
  c->Mass = Mass;
  c->Mass2 = Mass2;
  c->pexp = pexp;
  c->nthreads = nthreads > 0 ? nthreads : 1;
  size_t np = (size_t)Nmem * (Nmem + 1) / 2;
  /* The reference creates SymMatrix lattices as ones(...) (DefSimDataElements.jl:
     176).  No entry is read before it is written (init writes rows 1-2, each
     step writes its row before any later step reads it), so that initial
     content is not observable; zero pages are used instead. */
  c->bigbytes = np * c->V * sizeof(double);
  c->F = (double *)big_alloc(c->bigbytes);
  c->r = (double *)big_alloc(c->bigbytes);
  c->SigF = (double *)calloc((size_t)Nmem * c->V, sizeof(double));
  c->Sigr = (double *)calloc((size_t)Nmem * c->V, sizeof(double));
  if (pexp == 3) {
    c->Irx = (cplx *)calloc((size_t)Nmem * c->V, sizeof(cplx));
    c->IFx = (cplx *)calloc((size_t)Nmem * c->V, sizeof(cplx));
    if (!c->Irx || !c->IFx) {
      or_destroy(c);
      return NULL;
    }
  }
  c->kL2 = (double *)malloc(sizeof(double) * c->V);
  c->omega2 = (double *)malloc(sizeof(double) * c->V);
  if (!c->F || !c->r || !c->SigF || !c->Sigr || !c->kL2 || !c->omega2) {
    or_destroy(c);
    return NULL;
  }
  compute_kL2(c);
  reset_omega2(c);
  plan1d_init(&c->plan, Nx);
  c->tmin = 0;
  c->tmax = 0;
  return c;
}

void or_destroy(or_ctx *c) {
  if (!c) return;
  if (c->F) munmap(c->F, c->bigbytes);
  if (c->r) munmap(c->r, c->bigbytes);
  free(c->SigF);
  free(c->Sigr);
  free(c->Irx);
  free(c->IFx);
  free(c->kL2);
  free(c->omega2);
  if (c->plan.w) plan1d_free(&c->plan);
  free(c);
}

int or_get_kL2(or_ctx *c, double *out) {
  memcpy(out, c->kL2, sizeof(double) * c->V);
  return 0;
}

int or_set_kL2(or_ctx *c, const double *kL2) {
  memcpy(c->kL2, kL2, sizeof(double) * c->V);
  reset_omega2(c);
  return 0;
}

/* expandSimData!, SimDataStruct.jl:262-278 and expand!,
 * DefSimDataElements.jl:141-165 */
int or_expand(or_ctx *c) {
  if (c->tmax + 1 > c->N) {
    c->tmin += 1;
    c->tmax += 1;
  } else {
    c->tmin = 1;
    c->tmax += 1;
  }
  return 0;
}

int or_window(or_ctx *c, int *tmin, int *tmax) {
  *tmin = c->tmin;
  *tmax = c->tmax;
  return 0;
}

/* ------------------------------------------------------------------ */
/* 3D complex transform on the full grid                               */
/* ------------------------------------------------------------------ */
static void dft3(or_ctx *c, cplx *a, cplx *pen, cplx *tmp, int sign) {
  const int Nx = c->Nx;
  const int nx = Nx, ny = c->sdim >= 2 ? Nx : 1, nz = c->sdim >= 3 ? Nx : 1;
  /* x pencils */
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y) fft1d(&c->plan, a + ((size_t)z * ny + y) * nx, tmp, sign);
  if (ny > 1)
    for (int z = 0; z < nz; ++z)
      for (int x = 0; x < nx; ++x) {
        for (int y = 0; y < ny; ++y) pen[y] = a[((size_t)z * ny + y) * nx + x];
        fft1d(&c->plan, pen, tmp, sign);
        for (int y = 0; y < ny; ++y) a[((size_t)z * ny + y) * nx + x] = pen[y];
      }
  if (nz > 1)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        for (int z = 0; z < nz; ++z) pen[z] = a[((size_t)z * ny + y) * nx + x];
        fft1d(&c->plan, pen, tmp, sign);
        for (int z = 0; z < nz; ++z) a[((size_t)z * ny + y) * nx + x] = pen[z];
      }
}

int or_dft(or_ctx *c, double *data, int sign) {
  cplx *pen = (cplx *)malloc(sizeof(cplx) * c->Nx);
  cplx *tmp = (cplx *)malloc(sizeof(cplx) * c->Nx);
  dft3(c, (cplx *)data, pen, tmp, sign);
  free(pen);
  free(tmp);
  return 0;
}

/* ------------------------------------------------------------------ */
/* self energies                                                       */
/* ------------------------------------------------------------------ */
/* calcSelfEnergyattp!, CPUfull + LambdaNLOquantum, Evolution.jl:25-31:
 *   rx = real(rfft(r[tmone,tp] * thesign)),  Fx = real(rfft(F[tmone,tp]))
 *   Sigr[tp] = ivol^3 (-lambda^2 (N+2)/(6 N^2)) real(brfft(rx (Fx^2 - rx^2/12)))
 *   SigF[tp] = ivol^3 (-lambda^2 (N+2)/(18 N^2)) real(brfft(Fx (Fx^2 - 3/4 rx^2)))
 */
static void selfenergy_tp(or_ctx *c, int tmone, int tp, cplx *a, cplx *b,
                          cplx *pen, cplx *tmp) {
  const int V = c->V;
  const double ivol = 1.0 / (double)V;
  const double lam = c->lambda;
  const double ON = (double)c->ON;
  const double pre_r = ivol * ivol * ivol * (-(lam * lam) * (ON + 2.0) / (6.0 * ON * ON));
  const double pre_F = ivol * ivol * ivol * (-(lam * lam) * (ON + 2.0) / (18.0 * ON * ON));
  const double s = thesign(tmone, tp, c->N);
  const double *Fl = symptr(c, c->F, tmone, tp);
  const double *rl = symptr(c, c->r, tmone, tp);
  for (int i = 0; i < V; ++i) {
    a[i].re = Fl[i];
    a[i].im = 0.0;
    b[i].re = rl[i] * s;
    b[i].im = 0.0;
  }
  dft3(c, a, pen, tmp, -1);
  dft3(c, b, pen, tmp, -1);
  /* LambdaNLOclassical drops the rho^2 terms (Evolution.jl:107-120) */
  const double qr = c->pexp == 4 ? 0.0 : 1.0 / 12.0, qF = c->pexp == 4 ? 0.0 : 0.75;
  for (int i = 0; i < V; ++i) {
    double Fx = a[i].re, rx = b[i].re;
    double gr = rx * (Fx * Fx - qr * (rx * rx));
    double gF = Fx * (Fx * Fx - qF * (rx * rx));
    a[i].re = gF;
    a[i].im = 0.0;
    b[i].re = gr;
    b[i].im = 0.0;
  }
  dft3(c, a, pen, tmp, +1);
  dft3(c, b, pen, tmp, +1);
  double *SF = cycptr(c, c->SigF, tp);
  double *Sr = cycptr(c, c->Sigr, tp);
  for (int i = 0; i < V; ++i) {
    Sr[i] = pre_r * b[i].re;
    SF[i] = pre_F * a[i].re;
  }
}

static int ninv_selfenergies(or_ctx *c, int tmone, int simple);

/* calcSelfEnergies!, Evolution.jl:16-23 (threaded over tp); NinverseNLO:
 * Evolution.jl:144-154 */
int or_selfenergies(or_ctx *c, int tmone) {
  if (c->pexp == 0 || c->pexp == 2) return 0; /* LO: Evolution.jl:8-10 */
  if (c->pexp == 3) return ninv_selfenergies(c, tmone, 0);
  const int lo = c->tmin, hi = c->tmax - 1;
  if (!inwin(c, tmone)) return fail(c, "tmone %d outside window [%d,%d]", tmone, c->tmin, c->tmax);
#pragma omp parallel num_threads(c->nthreads)
  {
    cplx *a = (cplx *)malloc(sizeof(cplx) * c->V);
    cplx *b = (cplx *)malloc(sizeof(cplx) * c->V);
    cplx *pen = (cplx *)malloc(sizeof(cplx) * c->Nx);
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * c->Nx);
#pragma omp for schedule(dynamic, 1)
    for (int tp = lo; tp <= hi; ++tp) selfenergy_tp(c, tmone, tp, a, b, pen, tmp);
    free(a);
    free(b);
    free(pen);
    free(tmp);
  }
  return 0;
}

/* Brute-force twin of the FFT sunset: the momentum-space double convolution
 *   Sigma(p) = pre * V * sum_{q,k} X(q) Y(k) Z(p-q-k)
 * with X, Y, Z the inversion-symmetrised lattices (real(rfft(.)) only sees
 * the even part (X(p)+X(-p))/2).  O(V^3); for tiny lattices only.  This is
 * the role calcSelfEnergiessimple! (Evolution.jl:123-139) plays in the
 * reference's own fast-vs-simple test. */
int or_selfenergies_direct(or_ctx *c, int tmone) {
  if (c->pexp != 1 && c->pexp != 4) return fail(c, "direct self-energy: LambdaNLO only");
  const double qr = c->pexp == 4 ? 0.0 : 1.0 / 12.0, qF = c->pexp == 4 ? 0.0 : 0.75;
  const int V = c->V, Nx = c->Nx, sd = c->sdim;
  const double ivol = 1.0 / (double)V;
  const double lam = c->lambda, ON = (double)c->ON;
  const double pre_r = ivol * ivol * ivol * (-(lam * lam) * (ON + 2.0) / (6.0 * ON * ON));
  const double pre_F = ivol * ivol * ivol * (-(lam * lam) * (ON + 2.0) / (18.0 * ON * ON));
  double *Fe = (double *)malloc(sizeof(double) * V);
  double *re = (double *)malloc(sizeof(double) * V);
  int *neg = (int *)malloc(sizeof(int) * V);
  int(*co)[3] = malloc(sizeof(int[3]) * V);
  for (int i = 0; i < V; ++i) {
    int rem = i, m = 0, mul = 1;
    for (int d = 0; d < 3; ++d) {
      co[i][d] = 0;
      if (d < sd) {
        co[i][d] = rem % Nx;
        rem /= Nx;
        m += ((Nx - co[i][d]) % Nx) * mul;
        mul *= Nx;
      }
    }
    neg[i] = m;
  }
  for (int tp = c->tmin; tp <= c->tmax - 1; ++tp) {
    const double s = thesign(tmone, tp, c->N);
    const double *Fl = symptr(c, c->F, tmone, tp);
    const double *rl = symptr(c, c->r, tmone, tp);
    for (int i = 0; i < V; ++i) {
      Fe[i] = 0.5 * (Fl[i] + Fl[neg[i]]);
      re[i] = 0.5 * (rl[i] + rl[neg[i]]) * s;
    }
    double *SF = cycptr(c, c->SigF, tp);
    double *Sr = cycptr(c, c->Sigr, tp);
#pragma omp parallel for num_threads(c->nthreads) schedule(static)
    for (int p = 0; p < V; ++p) {
      double rFF = 0, rrr = 0, FFF = 0, Frr = 0;
      for (int q = 0; q < V; ++q)
        for (int k = 0; k < V; ++k) {
          int l = 0, mul = 1;
          for (int d = 0; d < sd; ++d) {
            int v = ((co[p][d] - co[q][d] - co[k][d]) % Nx + 2 * Nx) % Nx;
            l += v * mul;
            mul *= Nx;
          }
          rFF += re[q] * Fe[k] * Fe[l];
          rrr += re[q] * re[k] * re[l];
          FFF += Fe[q] * Fe[k] * Fe[l];
          Frr += Fe[q] * re[k] * re[l];
        }
      Sr[p] = pre_r * (double)V * (rFF - qr * rrr);
      SF[p] = pre_F * (double)V * (FFF - qF * Frr);
    }
  }
  free(Fe);
  free(re);
  free(neg);
  free(co);
  return 0;
}

/* ------------------------------------------------------------------ */
/* 1/N expansion at NLO (NinverseNLO), CPUfull                         */
/* ------------------------------------------------------------------ */
/* The reference works with rfft (x axis halved) and brfft.  Here every
 * lattice lives on the full grid: fwd(X) is the full complex DFT (sign -1),
 * whose restriction to the half grid is rfft(X); bwd(Y) = real part of the
 * unnormalised inverse DFT of the Hermitian extension Y, which is brfft(Y)
 * for the half-grid Y.  For a real lattice X, fwd(X)(-k) = conj(fwd(X)(k)),
 * and products of such arrays keep that symmetry, so the full-grid
 * arithmetic equals the reference's half-grid arithmetic entry by entry.
 * real(.) is taken exactly where the reference takes it.  Every lattice
 * the dynamics produces from the reference's initial states is inversion-
 * even (F(p) = F(-p)); for such lattices all transforms are real and the
 * fast, simple and momentum-space (CPUred2) forms coincide.  For lattices
 * with an odd part they do not even agree with each other (the fast form
 * takes real(Fx) real(rx), the simple one real(Fx rx)). */
typedef struct {
  cplx *a, *b, *g, *h, *Y, *pen, *tmp;
  double *RHS, *R2, *R3;
} ninv_ws;

static int thread_id(void) {
#ifdef _OPENMP
  return omp_get_thread_num();
#else
  return 0;
#endif
}

static void ninv_ws_alloc(or_ctx *c, ninv_ws *w) {
  const size_t V = (size_t)c->V;
  w->a = (cplx *)malloc(sizeof(cplx) * V);
  w->b = (cplx *)malloc(sizeof(cplx) * V);
  w->g = (cplx *)malloc(sizeof(cplx) * V);
  w->h = (cplx *)malloc(sizeof(cplx) * V);
  w->Y = (cplx *)malloc(sizeof(cplx) * V);
  w->pen = (cplx *)malloc(sizeof(cplx) * c->Nx);
  w->tmp = (cplx *)malloc(sizeof(cplx) * c->Nx);
  w->RHS = (double *)malloc(sizeof(double) * V);
  w->R2 = (double *)malloc(sizeof(double) * V);
  w->R3 = (double *)malloc(sizeof(double) * V);
}

static void ninv_ws_free(ninv_ws *w) {
  free(w->a);
  free(w->b);
  free(w->g);
  free(w->h);
  free(w->Y);
  free(w->pen);
  free(w->tmp);
  free(w->RHS);
  free(w->R2);
  free(w->R3);
}

/* out = rfft(x * s) on the full grid */
static void nfwd(or_ctx *c, ninv_ws *w, const double *x, double s, cplx *out) {
#pragma omp atomic
  c->ntransforms += 1;
  for (int i = 0; i < c->V; ++i) {
    out[i].re = x[i] * s;
    out[i].im = 0.0;
  }
  dft3(c, out, w->pen, w->tmp, -1);
}

/* out = brfft(Y) (real lattice); Y is preserved */
static void nbwd(or_ctx *c, ninv_ws *w, const cplx *Y, double *out) {
#pragma omp atomic
  c->ntransforms += 1;
  for (int i = 0; i < c->V; ++i) w->h[i] = Y[i];
  dft3(c, w->h, w->pen, w->tmp, +1);
  for (int i = 0; i < c->V; ++i) out[i] = w->h[i].re;
}

static cplx *ircyc(or_ctx *c, cplx *base, int i) { return base + (size_t)(mod1(i, c->N) - 1) * (size_t)c->V; }

/* calcIrs! CPUfull, Evolution.jl:170-206 (tp counts down: Ir(tp) needs
 * Ir(z) for z > tp) */
/* w: nt workspaces, one per thread */
static void ninv_Irs(or_ctx *c, ninv_ws *w, int nt, int tmone) {
  const int V = c->V, N = c->N;
  const double ivol = 1.0 / (double)V;
  const double f1 = (c->lambda / 3.0) * ivol * ivol;
  const double f2 = c->dt * (c->lambda / 3.0) * ivol;
  const double f3 = ivol * ivol;
  cplx *It = ircyc(c, c->Irx, tmone);
  for (int i = 0; i < V; ++i) It[i].re = It[i].im = 0.0;
  for (int tp = tmone - 1; tp >= c->tmin; --tp) {
    nfwd(c, w, symptr(c, c->F, tmone, tp), 1.0, w->a);
    nfwd(c, w, symptr(c, c->r, tmone, tp), 1.0, w->b);
    const double s = thesign(tmone, tp, N);
    cplx *I = ircyc(c, c->Irx, tp);
    for (int i = 0; i < V; ++i) {
      I[i].re = s * f1 * w->a[i].re * w->b[i].re;
      I[i].im = 0.0;
    }
#pragma omp parallel num_threads(nt)
    {
      ninv_ws *u = &w[thread_id()];
      for (int i = 0; i < V; ++i) u->RHS[i] = 0.0;
#pragma omp for schedule(static)
      for (int z = tp + 1; z <= tmone - 1; ++z) {
        nfwd(c, u, symptr(c, c->F, z, tp), 1.0, u->a);
        nfwd(c, u, symptr(c, c->r, z, tp), 1.0, u->b);
        const double sz = thesign(z, tp, N);
        for (int i = 0; i < V; ++i) {
          u->g[i].re = u->a[i].re * u->b[i].re;
          u->g[i].im = 0.0;
        }
        nbwd(c, u, u->g, u->R2);
        nbwd(c, u, ircyc(c, c->Irx, z), u->R3);
        for (int i = 0; i < V; ++i) u->RHS[i] -= f3 * sz * u->R2[i] * u->R3[i];
      }
    }
    for (int k = 1; k < nt; ++k)
      for (int i = 0; i < V; ++i) w->RHS[i] += w[k].RHS[i];
    nfwd(c, w, w->RHS, 1.0, w->g);
    for (int i = 0; i < V; ++i) I[i].re += f2 * w->g[i].re;
  }
}

/* calcIrssimple!, Evolution.jl:294-314 */
static void ninv_Irs_simple(or_ctx *c, ninv_ws *w, int tmone) {
  const int V = c->V, N = c->N;
  const double ivol = 1.0 / (double)V;
  const double lam = c->lambda;
  cplx *It = ircyc(c, c->Irx, tmone);
  for (int i = 0; i < V; ++i) It[i].re = It[i].im = 0.0;
  for (int tp = tmone - 1; tp >= c->tmin; --tp) {
    nfwd(c, w, symptr(c, c->F, tmone, tp), 1.0, w->a);
    nfwd(c, w, symptr(c, c->r, tmone, tp), thesign(tmone, tp, N), w->b);
    cplx *I = ircyc(c, c->Irx, tp);
    for (int i = 0; i < V; ++i) {
      /* real((lambda/3) ivol^2 rfft(F) .* rfft(r s)) */
      I[i].re = (lam / 3.0) * ivol * ivol * (w->a[i].re * w->b[i].re - w->a[i].im * w->b[i].im);
      I[i].im = 0.0;
    }
    for (int z = tp + 1; z <= tmone - 1; ++z) {
      nfwd(c, w, symptr(c, c->F, z, tp), 1.0, w->a);
      nfwd(c, w, symptr(c, c->r, z, tp), thesign(z, tp, N), w->b);
      for (int i = 0; i < V; ++i) {
        w->g[i].re = ivol * ivol * (w->a[i].re * w->b[i].re - w->a[i].im * w->b[i].im); /* Bx */
        w->g[i].im = 0.0;
      }
      nbwd(c, w, ircyc(c, c->Irx, z), w->R3);
      nbwd(c, w, w->g, w->R2);
      for (int i = 0; i < V; ++i) w->RHS[i] = w->R3[i] * w->R2[i];
      nfwd(c, w, w->RHS, 1.0, w->g);
      for (int i = 0; i < V; ++i) I[i].re -= c->dt * (lam / 3.0) * ivol * w->g[i].re;
    }
  }
}

/* F2 = fwd(F)^2 - 1/4 fwd(r)^2 (complex), scaled by sc */
static void ninv_F2r2(or_ctx *c, ninv_ws *w, int z, int tp, double sc, cplx *out) {
  nfwd(c, w, symptr(c, c->F, z, tp), 1.0, w->a);
  nfwd(c, w, symptr(c, c->r, z, tp), 1.0, w->b);
  for (int i = 0; i < c->V; ++i) {
    const double Fr = w->a[i].re * sc, Fi = w->a[i].im * sc, rr = w->b[i].re * sc, ri = w->b[i].im * sc;
    out[i].re = (Fr * Fr - Fi * Fi) - 0.25 * (rr * rr - ri * ri);
    out[i].im = 2.0 * Fr * Fi - 0.25 * 2.0 * rr * ri;
  }
}

/* fwd(F) fwd(r) s (complex), scaled by sc^2 */
static void ninv_Fr(or_ctx *c, ninv_ws *w, int z, int tp, double sc, double s, cplx *out) {
  nfwd(c, w, symptr(c, c->F, z, tp), 1.0, w->a);
  nfwd(c, w, symptr(c, c->r, z, tp), 1.0, w->b);
  for (int i = 0; i < c->V; ++i) {
    const double Fr = w->a[i].re * sc, Fi = w->a[i].im * sc, rr = w->b[i].re * sc, ri = w->b[i].im * sc;
    out[i].re = s * (Fr * rr - Fi * ri);
    out[i].im = s * (Fr * ri + Fi * rr);
  }
}

/* calcIFs! CPUfull, Evolution.jl:354-436 (tp counts up: IF(tp) needs IF(z)
 * for z < tp).  IFx[tp] += ivol * Fx is complex in the reference; its
 * imaginary part is kept. */
static void ninv_IFs(or_ctx *c, ninv_ws *w, int nt, int tmone) {
  const int V = c->V, N = c->N, tmin = c->tmin;
  const double ivol = 1.0 / (double)V;
  const double lam = c->lambda, dt = c->dt;
  const double f1 = lam / 6.0 * ivol * ivol, f2 = lam / 6.0 * dt, f3 = lam / 3.0 * dt;
  for (int tp = tmin; tp <= tmone; ++tp) {
    cplx *IF = ircyc(c, c->IFx, tp);
    cplx *Y = w->Y;
    ninv_F2r2(c, w, tmone, tp, 1.0, Y);
    for (int i = 0; i < V; ++i) {
      IF[i].re = f1 * Y[i].re; /* real(factor1 * Fx) */
      IF[i].im = 0.0;
    }
    const int beg = tp == tmin ? tmin + 1 : tp;
#pragma omp parallel num_threads(nt)
    {
      ninv_ws *u = &w[thread_id()];
      cplx *Yu = u->Y;
      for (int i = 0; i < V; ++i) u->RHS[i] = 0.0;
      /* first integral, z = tmin+1 .. tp-1: Ir part (nr 1) and IF part (nr 2) */
#pragma omp for schedule(static)
      for (int z = tmin + 1; z <= tp - 1; ++z) {
        nbwd(c, u, ircyc(c, c->Irx, z), u->R3);
        ninv_F2r2(c, u, z, tp, ivol, Yu);
        nbwd(c, u, Yu, u->R2);
        for (int i = 0; i < V; ++i) u->RHS[i] -= f2 * u->R3[i] * u->R2[i];
        ninv_Fr(c, u, z, tp, ivol, 1.0, Yu);
        const double s = thesign(z, tp, N);
        nbwd(c, u, Yu, u->R3);
        nbwd(c, u, ircyc(c, c->IFx, z), u->R2);
        for (int i = 0; i < V; ++i) u->RHS[i] += f3 * s * u->R2[i] * u->R3[i];
      }
      /* second integral, z = max(tp, tmin+1) .. tmone-1: Ir part only */
#pragma omp for schedule(static)
      for (int z = beg; z <= tmone - 1; ++z) {
        ninv_F2r2(c, u, z, tp, ivol, Yu);
        nbwd(c, u, Yu, u->R3);
        nbwd(c, u, ircyc(c, c->Irx, z), u->R2);
        for (int i = 0; i < V; ++i) u->RHS[i] -= f2 * u->R2[i] * u->R3[i];
      }
    }
    for (int k = 1; k < nt; ++k)
      for (int i = 0; i < V; ++i) w->RHS[i] += w[k].RHS[i];
    /* border z = tmin, trapezoid weight 1/2 */
    ninv_F2r2(c, w, tmin, tp, 1.0, Y);
    nbwd(c, w, Y, w->R3);
    for (int i = 0; i < V; ++i) w->R3[i] *= ivol * ivol;
    nbwd(c, w, ircyc(c, c->Irx, tmin), w->R2);
    for (int i = 0; i < V; ++i) w->RHS[i] -= 0.5 * dt * lam / 6.0 * w->R2[i] * w->R3[i];
    if (tp > tmin) {
      ninv_Fr(c, w, tmin, tp, 1.0, thesign(tmin, tp, N), Y);
      nbwd(c, w, Y, w->R3);
      nbwd(c, w, ircyc(c, c->IFx, tmin), w->R2);
      for (int i = 0; i < V; ++i) w->RHS[i] += 0.5 * dt * lam / 3.0 * ivol * ivol * w->R2[i] * w->R3[i];
    }
    nfwd(c, w, w->RHS, 1.0, w->g);
    for (int i = 0; i < V; ++i) {
      IF[i].re += ivol * w->g[i].re;
      IF[i].im += ivol * w->g[i].im;
    }
  }
}

/* calcIFssimple!, Evolution.jl:599-624 */
static void ninv_IFs_simple(or_ctx *c, ninv_ws *w, int tmone) {
  const int V = c->V, N = c->N, tmin = c->tmin;
  const double ivol = 1.0 / (double)V;
  const double lam = c->lambda, dt = c->dt;
  cplx *Y = (cplx *)malloc(sizeof(cplx) * V);
  for (int tp = tmin; tp <= tmone; ++tp) {
    cplx *IF = ircyc(c, c->IFx, tp);
    ninv_F2r2(c, w, tmone, tp, 1.0, Y);
    for (int i = 0; i < V; ++i) {
      IF[i].re = (lam / 6.0) * ivol * ivol * Y[i].re;
      IF[i].im = 0.0;
    }
    for (int z = tmin; z <= tmone - 1; ++z) {
      const double wz = z == tmin ? 0.5 * dt : dt;
      nbwd(c, w, ircyc(c, c->Irx, z), w->R3);
      ninv_F2r2(c, w, z, tp, 1.0, Y);
      for (int i = 0; i < V; ++i) {
        Y[i].re *= ivol * ivol;
        Y[i].im *= ivol * ivol;
      }
      nbwd(c, w, Y, w->R2);
      for (int i = 0; i < V; ++i) w->RHS[i] = w->R3[i] * w->R2[i];
      nfwd(c, w, w->RHS, 1.0, w->g);
      for (int i = 0; i < V; ++i) IF[i].re -= wz * (lam / 6.0) * ivol * w->g[i].re;
    }
    for (int z = tmin; z <= tp - 1; ++z) {
      const double wz = z == tmin ? 0.5 * dt : dt;
      nbwd(c, w, ircyc(c, c->IFx, z), w->R3);
      ninv_Fr(c, w, z, tp, 1.0, thesign(z, tp, N), Y);
      for (int i = 0; i < V; ++i) {
        Y[i].re *= ivol * ivol;
        Y[i].im *= ivol * ivol;
      }
      nbwd(c, w, Y, w->R2);
      for (int i = 0; i < V; ++i) w->RHS[i] = w->R3[i] * w->R2[i];
      nfwd(c, w, w->RHS, 1.0, w->g);
      for (int i = 0; i < V; ++i) IF[i].re += wz * (lam / 3.0) * ivol * w->g[i].re;
    }
  }
  free(Y);
}

/* calcSelfEnergyattp! CPUfull NinverseNLO, Evolution.jl:683-698 (fast)
 * and calcSelfEnergyattpsimple!, :737-742: both evaluate
 *   Sigr = -lambda/(3N) ivol brfft(Fx Irx + rx IFx)
 *   SigF = -lambda/(3N) ivol brfft(Fx IFx - 1/4 rx Irx)
 * with Fx = rfft(F(tmone,tp)), rx = rfft(r(tmone,tp)) * thesign. */
static void ninv_sigma_tp(or_ctx *c, ninv_ws *w, int tmone, int tp) {
  const int V = c->V;
  const double ivol = 1.0 / (double)V;
  const double pre = -c->lambda / (3.0 * (double)c->ON) * ivol;
  nfwd(c, w, symptr(c, c->F, tmone, tp), 1.0, w->a);
  nfwd(c, w, symptr(c, c->r, tmone, tp), thesign(tmone, tp, c->N), w->b);
  const cplx *Ir = ircyc(c, c->Irx, tp), *IF = ircyc(c, c->IFx, tp);
  for (int i = 0; i < V; ++i) {
    const cplx A = w->a[i], B = w->b[i], R = Ir[i], Q = IF[i];
    w->g[i].re = (A.re * R.re - A.im * R.im) + (B.re * Q.re - B.im * Q.im);
    w->g[i].im = (A.re * R.im + A.im * R.re) + (B.re * Q.im + B.im * Q.re);
  }
  nbwd(c, w, w->g, cycptr(c, c->Sigr, tp));
  for (int i = 0; i < V; ++i) {
    const cplx A = w->a[i], B = w->b[i], R = Ir[i], Q = IF[i];
    w->g[i].re = (A.re * Q.re - A.im * Q.im) - 0.25 * (B.re * R.re - B.im * R.im);
    w->g[i].im = (A.re * Q.im + A.im * Q.re) - 0.25 * (B.re * R.im + B.im * R.re);
  }
  nbwd(c, w, w->g, cycptr(c, c->SigF, tp));
  double *Sr = cycptr(c, c->Sigr, tp), *SF = cycptr(c, c->SigF, tp);
  for (int i = 0; i < V; ++i) {
    Sr[i] *= pre;
    SF[i] *= pre;
  }
}

/* calcSelfEnergies! NinverseNLO, Evolution.jl:144-154: Ir and IF for all
 * tp, then the self-energies for tp in [tmin, tmone] */
static int ninv_selfenergies(or_ctx *c, int tmone, int simple) {
  if (!inwin(c, tmone)) return fail(c, "tmone %d outside window [%d,%d]", tmone, c->tmin, c->tmax);
  const int nt = simple ? 1 : c->nthreads;
  ninv_ws *w = (ninv_ws *)malloc(sizeof(ninv_ws) * nt);
  for (int k = 0; k < nt; ++k) ninv_ws_alloc(c, &w[k]);
  if (simple) {
    ninv_Irs_simple(c, w, tmone);
    ninv_IFs_simple(c, w, tmone);
  } else {
    ninv_Irs(c, w, nt, tmone);
    ninv_IFs(c, w, nt, tmone);
  }
  const int lo = c->tmin, hi = c->tmax - 1;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
  for (int tp = lo; tp <= hi; ++tp) ninv_sigma_tp(c, &w[thread_id()], tmone, tp);
  for (int k = 0; k < nt; ++k) ninv_ws_free(&w[k]);
  free(w);
  return 0;
}

int or_ninv_resum(or_ctx *c, int tmone, int simple) {
  if (c->pexp != 3) return fail(c, "resummation: NinverseNLO only");
  if (!inwin(c, tmone)) return fail(c, "tmone %d outside window [%d,%d]", tmone, c->tmin, c->tmax);
  const int nt = simple ? 1 : c->nthreads;
  ninv_ws *w = (ninv_ws *)malloc(sizeof(ninv_ws) * nt);
  for (int k = 0; k < nt; ++k) ninv_ws_alloc(c, &w[k]);
  if (simple) {
    ninv_Irs_simple(c, w, tmone);
    ninv_IFs_simple(c, w, tmone);
  } else {
    ninv_Irs(c, w, nt, tmone);
    ninv_IFs(c, w, nt, tmone);
  }
  for (int k = 0; k < nt; ++k) ninv_ws_free(&w[k]);
  free(w);
  return 0;
}

int or_ninv_selfenergies(or_ctx *c, int tmone, int simple) {
  if (c->pexp != 3) return fail(c, "1/N self-energies: NinverseNLO only");
  return ninv_selfenergies(c, tmone, simple);
}

long long or_ninv_transforms(or_ctx *c) { return c->ntransforms; }

/* Irx / IFx at the slot of time tp (which = 0 IFx, 1 Irx), as V complex
 * values (re, im interleaved) */
int or_get_ninv(or_ctx *c, int which, int tp, double *out) {
  if (c->pexp != 3) return fail(c, "no 1/N resummation in this context");
  if (!inwin(c, tp)) return fail(c, "ninv: %d outside window", tp);
  memcpy(out, ircyc(c, which ? c->Irx : c->IFx, tp), sizeof(cplx) * c->V);
  return 0;
}

/* Hartreeprefactor, Initialization.jl:245-247: lambda (N+2)/(6N) for the
 * lambda expansions and NinverseNLO, lambda/6 for NinverseLO */
static double hartreeprefactor(const or_ctx *c) {
  if (c->pexp == 2) return c->lambda * (1.0 / 6.0);
  return c->lambda * (double)(c->ON + 2) / (6.0 * (double)c->ON);
}

/* getHartreeMass2 CPUfull, Initialization.jl:207-218 */
double or_hartree(or_ctx *c, int i, int j) {
  const double *Fl = symptr(c, c->F, i, j);
  double h = 0.0;
  for (int k = 0; k < c->V; ++k) h += Fl[k];
  h *= hartreeprefactor(c) * (1.0 / (double)c->V);
  h += c->Mass2;
  return h;
}

int or_set_hartree(or_ctx *c, double m2) {
  c->hartreemass2 = m2;
  return 0;
}
double or_get_hartree(or_ctx *c) { return c->hartreemass2; }

/* setomega2values!, Evolution.jl:918-921 */
static void setomega2(or_ctx *c) {
  for (int i = 0; i < c->V; ++i) {
    c->omega2[i] = c->kL2[i] + c->hartreemass2;
    c->omega2[i] *= c->dt * c->dt;
  }
}

/* getRHS_r! CPUfull NLO, Evolution.jl:846-854 */
static void rhs_r(or_ctx *c, double *Mem, int t, int tp) {
  const int V = c->V;
  for (int i = 0; i < V; ++i) Mem[i] = 0.0;
  if (c->pexp == 0 || c->pexp == 2) return; /* LO: Evolution.jl:842-844 */
  for (int z = tp + 1; z <= t - 2; ++z) {
    const double *S = cycptr(c, c->Sigr, z);
    const double *R = symptr(c, c->r, z, tp);
    const double s = thesign(z, tp, c->N);
    for (int i = 0; i < V; ++i) Mem[i] -= S[i] * R[i] * s;
  }
}

/* getRHS_F! CPUfull NLO, Evolution.jl:766-782 */
static void rhs_F(or_ctx *c, double *Mem, int t, int tp) {
  const int V = c->V;
  const int tpmin = c->tmin;
  if (c->pexp == 0 || c->pexp == 2) { /* LO: Evolution.jl:763-765 */
    for (int i = 0; i < V; ++i) Mem[i] = 0.0;
    return;
  }
  {
    const double *S = cycptr(c, c->Sigr, tpmin);
    const double *Fl = symptr(c, c->F, tpmin, tp);
    for (int i = 0; i < V; ++i) {
      Mem[i] = S[i] * Fl[i];
      Mem[i] *= -0.5;
    }
  }
  for (int z = tpmin + 1; z <= t - 2; ++z) {
    const double *S = cycptr(c, c->Sigr, z);
    const double *Fl = symptr(c, c->F, z, tp);
    for (int i = 0; i < V; ++i) Mem[i] -= S[i] * Fl[i];
  }
  {
    const double *S = cycptr(c, c->SigF, tpmin);
    const double *R = symptr(c, c->r, tpmin, tp);
    const double s = thesign(tpmin, tp, c->N);
    for (int i = 0; i < V; ++i) Mem[i] += 0.5 * S[i] * R[i] * s;
  }
  for (int z = tpmin + 1; z <= tp - 1; ++z) {
    const double *S = cycptr(c, c->SigF, z);
    const double *R = symptr(c, c->r, z, tp);
    const double s = thesign(z, tp, c->N);
    for (int i = 0; i < V; ++i) Mem[i] += S[i] * R[i] * s;
  }
}

/* evolve_r!, Evolution.jl:890-905 */
static void evolve_r(or_ctx *c, double *RHS, int t, int tp) {
  const int V = c->V, N = c->N;
  double *out = symptr(c, c->r, t, tp);
  if (tp == t - 1) {
    const double v = thesign(t, tp, N) * c->dt;
    for (int i = 0; i < V; ++i) out[i] = v;
  } else if (tp == t) {
    for (int i = 0; i < V; ++i) out[i] = 0.0;
  } else {
    const double dt3 = c->dt * c->dt * c->dt;
    const double *r1 = symptr(c, c->r, t - 1, tp);
    const double *r2 = symptr(c, c->r, t - 2, tp);
    const double s1 = thesign(t - 1, tp, N), s2 = thesign(t - 2, tp, N);
    const double s0 = thesign(t, tp, N);
    for (int i = 0; i < V; ++i) {
      RHS[i] *= dt3;
      out[i] = (2.0 - c->omega2[i]) * r1[i] * s1 - r2[i] * s2 + RHS[i];
      out[i] *= s0;
    }
  }
}

/* evolve_F!, Evolution.jl:908-915 */
static void evolve_F(or_ctx *c, double *RHS, int t, int tp) {
  const int V = c->V;
  const double dt3 = c->dt * c->dt * c->dt;
  double *out = symptr(c, c->F, t, tp);
  const double *F1 = symptr(c, c->F, t - 1, tp);
  const double *F2 = symptr(c, c->F, t - 2, tp);
  for (int i = 0; i < V; ++i) {
    RHS[i] *= dt3;
    out[i] = (2.0 - c->omega2[i]) * F1[i] - F2[i] + RHS[i];
  }
}

/* evolve!, Evolution.jl:924-981 */
int or_evolve(or_ctx *c, int t) {
  if (t != c->tmax + 1) return fail(c, "evolve: t=%d but window is [%d,%d]", t, c->tmin, c->tmax);
  if (t < 3) return fail(c, "evolve: t=%d < 3", t);
  or_expand(c);
  /* calcSelfEnergies! at tmone (Evolution.jl:942) */
  or_selfenergies(c, t - 1);
  /* hartree mass and omega2 (:945-947) */
  c->hartreemass2 = or_hartree(c, t - 1, t - 1);
  setomega2(c);
  /* off-diagonal rows (:951-967) */
  const int lo = c->tmin, hi = c->tmax - 1;
#pragma omp parallel num_threads(c->nthreads)
  {
    double *RHS = (double *)malloc(sizeof(double) * c->V);
#pragma omp for schedule(dynamic, 1)
    for (int tp = lo; tp <= hi; ++tp) {
      rhs_r(c, RHS, t, tp);
      evolve_r(c, RHS, t, tp);
      rhs_F(c, RHS, t, tp);
      evolve_F(c, RHS, t, tp);
    }
    free(RHS);
  }
  /* diagonal (:971-974) */
  double *RHS = (double *)malloc(sizeof(double) * c->V);
  evolve_r(c, RHS, t, t);
  rhs_F(c, RHS, t, t);
  evolve_F(c, RHS, t, t);
  free(RHS);
  return 0;
}

/* initSimData! CPUfull, Initialization.jl:8-29 */
int or_init_particles(or_ctx *c, const double *n, double initmass) {
  if (c->tmax != 0) return fail(c, "init: context already initialised");
  or_expand(c);
  or_expand(c);
  const int N = c->N;
  double *F11 = symptr(c, c->F, 1, 1), *F21 = symptr(c, c->F, 2, 1), *F22 = symptr(c, c->F, 2, 2);
  double *r11 = symptr(c, c->r, 1, 1), *r21 = symptr(c, c->r, 2, 1), *r22 = symptr(c, c->r, 2, 2);
  for (int i = 0; i < c->V; ++i) {
    double omega = sqrt(c->kL2[i] + initmass * initmass);
    F11[i] = (1.0 / omega) * (n[i] + 0.5);
    F21[i] = (1.0 / omega) * (n[i] + 0.5);
    F22[i] = (omega * (n[i] + 0.5) * (c->dt * c->dt) + (1.0 / omega) * (n[i] + 0.5));
    r11[i] = 0.0;
    r21[i] = c->dt * thesign(2, 1, N);
    r22[i] = 0.0;
  }
  c->hartreemass2 = or_hartree(c, 2, 2);
  return 0;
}

/* initSimData! Gauss, Initialization.jl:95-112 (rows given) */
int or_init_rows(or_ctx *c, const double *F11in, const double *F21in, const double *F22in) {
  if (c->tmax != 0) return fail(c, "init: context already initialised");
  or_expand(c);
  or_expand(c);
  const int N = c->N;
  double *F11 = symptr(c, c->F, 1, 1), *F21 = symptr(c, c->F, 2, 1), *F22 = symptr(c, c->F, 2, 2);
  double *r11 = symptr(c, c->r, 1, 1), *r21 = symptr(c, c->r, 2, 1), *r22 = symptr(c, c->r, 2, 2);
  for (int i = 0; i < c->V; ++i) {
    F11[i] = F11in[i];
    F21[i] = F21in[i];
    F22[i] = F22in[i];
    r11[i] = 0.0;
    r21[i] = c->dt * thesign(2, 1, N);
    r22[i] = 0.0;
  }
  c->hartreemass2 = or_hartree(c, 2, 2);
  return 0;
}

int or_get(or_ctx *c, int which, int i, int j, double *out) {
  if (!inwin(c, i) || !inwin(c, j)) return fail(c, "get: (%d,%d) outside window [%d,%d]", i, j, c->tmin, c->tmax);
  const double *p = symptr(c, which ? c->r : c->F, i, j);
  const double s = which ? thesign(i, j, c->N) : 1.0;
  for (int k = 0; k < c->V; ++k) out[k] = p[k] * s;
  return 0;
}

int or_set(or_ctx *c, int which, int i, int j, const double *in) {
  if (!inwin(c, i) || !inwin(c, j)) return fail(c, "set: (%d,%d) outside window [%d,%d]", i, j, c->tmin, c->tmax);
  double *p = symptr(c, which ? c->r : c->F, i, j);
  const double s = which ? thesign(i, j, c->N) : 1.0;
  for (int k = 0; k < c->V; ++k) p[k] = in[k] * s;
  return 0;
}

int or_get_sigma(or_ctx *c, int which, int tp, double *out) {
  if (!inwin(c, tp)) return fail(c, "sigma: %d outside window", tp);
  memcpy(out, cycptr(c, which ? c->Sigr : c->SigF, tp), sizeof(double) * c->V);
  return 0;
}

int or_set_sigma(or_ctx *c, int which, int tp, const double *in) {
  if (!inwin(c, tp)) return fail(c, "sigma: %d outside window", tp);
  memcpy(cycptr(c, which ? c->Sigr : c->SigF, tp), in, sizeof(double) * c->V);
  return 0;
}

/* createMeasurement + getE_kin/getE_pot/getE_mem (CPUfull, lattice),
 * Measurement.jl:31-99, 135-174, evaluated with every site its own level */
int or_measure(or_ctx *c, int t, double *scal, double *Fo, double *ddFo,
               double *ro, double *no, double *omo, double *co, int *nstab) {
  if (!inwin(c, t) || !inwin(c, t - 1)) return fail(c, "measure: t=%d outside window", t);
  const int V = c->V, N = c->N;
  const double dt = c->dt;
  const double *Ftt = symptr(c, c->F, t, t);
  const double *Fmt = symptr(c, c->F, t - 1, t);
  const double *Fmm = symptr(c, c->F, t - 1, t - 1);
  const double *rmm = symptr(c, c->r, t - 1, t - 1);
  const double smm = thesign(t - 1, t - 1, N);
  double Ekin = 0.0, hsum = 0.0, Epot = 0.0;
  int bad = 0;
  for (int i = 0; i < V; ++i) {
    double ddF = (Ftt[i] - 2.0 * Fmt[i] + Fmm[i]) / (dt * dt);
    double F = Fmm[i];
    double om = (ddF / F) > 0 ? sqrt(ddF / F) : 0.0;
    double cc;
    if (0.25 * (dt * dt) * (om * om) < 1.0)
      cc = sqrt(1.0 - 0.25 * (dt * dt) * (om * om));
    else {
      cc = 1.0;
      bad++;
    }
    double n = (ddF * F > 0) ? (cc * sqrt(ddF * F)) - 0.5 : 0.0;
    if (Fo) Fo[i] = F;
    if (ddFo) ddFo[i] = ddF;
    if (ro) ro[i] = rmm[i] * smm;
    if (no) no[i] = n;
    if (omo) omo[i] = om;
    if (co) co[i] = cc;
    Ekin += 0.5 * ddF;
    hsum += F;
  }
  const double H = hartreeprefactor(c);
  const double ivol = 1.0 / (double)V;
  for (int i = 0; i < V; ++i)
    Epot += 0.5 * (c->Mass2 + c->kL2[i] + 0.5 * ivol * H * hsum) * Fmm[i];
  double Emem = 0.0;
  if (t != 2) {
    const int tmin = c->tmin;
    {
      const double *S = cycptr(c, c->Sigr, tmin), *SF = cycptr(c, c->SigF, tmin);
      const double *Fl = symptr(c, c->F, tmin, t - 1), *R = symptr(c, c->r, tmin, t - 1);
      const double s = thesign(tmin, t - 1, N);
      double acc = 0.0;
      for (int i = 0; i < V; ++i) acc += S[i] * Fl[i] - SF[i] * R[i] * s;
      Emem = 0.5 * 0.25 * dt * acc;
    }
    if (t > 3) {
      double tot = 0.0;
      for (int tp = tmin + 1; tp <= t - 2; ++tp) {
        const double *S = cycptr(c, c->Sigr, tp), *SF = cycptr(c, c->SigF, tp);
        const double *Fl = symptr(c, c->F, tp, t - 1), *R = symptr(c, c->r, tp, t - 1);
        const double s = thesign(tp, t - 1, N);
        double acc = 0.0;
        for (int i = 0; i < V; ++i) acc += S[i] * Fl[i] - SF[i] * R[i] * s;
        tot += acc;
      }
      Emem += 0.25 * dt * tot;
    }
  }
  if (scal) {
    scal[0] = (double)(t - 1) * dt * c->Mass;
    scal[2] = Epot;
    scal[3] = Ekin;
    scal[4] = Emem;
    scal[1] = Ekin + Epot + Emem;
    scal[5] = c->hartreemass2;
  }
  if (nstab) *nstab = bad;
  return 0;
}

/* Benchmark set-up, as the reference's benchmark harness fills the window
 * without physics (TwoPIScalar_Bechmark.jl:62-69): a full window
 * [tmax-N+1, tmax] of free propagators F = (n+1/2)/w cos(w(a-b)dt),
 * rho(a,b) = sin(w(a-b)dt)/w, w = sqrt(kL2 + m^2).  n may be NULL. */
int or_fill_free_window(or_ctx *c, int tmax, const double *n, double mass) {
  if (tmax < c->N) return fail(c, "fill: tmax %d < N %d", tmax, c->N);
  c->tmin = tmax - c->N + 1;
  c->tmax = tmax;
  const int V = c->V, N = c->N;
  const int tmin = c->tmin;
  /* the free propagators depend on the lag a-b only: tabulate N lags */
  double *Ft = (double *)malloc(sizeof(double) * (size_t)N * V);
  double *rt = (double *)malloc(sizeof(double) * (size_t)N * V);
  if (!Ft || !rt) {
    free(Ft);
    free(rt);
    return fail(c, "fill: out of memory");
  }
#pragma omp parallel for num_threads(c->nthreads) schedule(static)
  for (int d = 0; d < N; ++d)
    for (int i = 0; i < V; ++i) {
      const double w = sqrt(c->kL2[i] + mass * mass);
      const double nn = n ? n[i] : 0.0;
      const double x = (double)d * c->dt;
      Ft[(size_t)d * V + i] = (nn + 0.5) / w * cos(w * x);
      rt[(size_t)d * V + i] = sin(w * x) / w;
    }
#pragma omp parallel for num_threads(c->nthreads) schedule(dynamic, 1)
  for (int a = tmin; a <= tmax; ++a)
    for (int b = tmin; b <= a; ++b) {
      double *Fl = symptr(c, c->F, a, b), *rl = symptr(c, c->r, a, b);
      const double s = thesign(a, b, N);
      const double *Fs = Ft + (size_t)(a - b) * V, *rs = rt + (size_t)(a - b) * V;
      memcpy(Fl, Fs, sizeof(double) * V);
      for (int i = 0; i < V; ++i) rl[i] = rs[i] * s;
    }
  free(Ft);
  free(rt);
  c->hartreemass2 = or_hartree(c, tmax, tmax);
  return 0;
}
