// kb_api.hip -- host side of libqftkb: the extern "C" boundary declared in
// include/qftkb.h.  Owns device memory, window bookkeeping (the reference's
// simdata.indices, src/DefSimDataElements.jl:141-165) and kernel launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/qftkb.h"
#include "kb_internal.h"

#ifndef KB_MSPLIT_MAX
#define KB_MSPLIT_MAX 8
#endif
// largest orbit count whose self-energy runs as orbit products (the matrix
// takes orbits^2 doubles: 134 MB at the limit)
// lattice limits: the pencil passes hold two buffers of Nx complex values per
// pencil in LDS (64 KiB at Nx = 2048, one pencil per workgroup); with at most
// 2^24 sites a launch of one workgroup per 8-site block stays below 2^32
// work-items at 1024 threads per workgroup
constexpr int KB_MAX_NX = 2048;
constexpr long KB_MAX_SITES = 1L << 24;

#ifndef KB_ORBIT_PRODUCT_MAX
#define KB_ORBIT_PRODUCT_MAX 4096
#endif

using namespace qftkb;
using namespace qftkb_host;

// host helpers: internal to the library, not part of its dynamic symbol table
#pragma GCC visibility push(hidden)
namespace qftkb_host {

int seterr(kb_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

inline int slot_of(const kb_ctx* c, int t) { return (t - 1) % c->N; }
inline bool inwin(const kb_ctx* c, int t) { return c->st.tmin <= t && t <= c->st.tmax; }

void plan_fft(FFTPlan& pl, int n) {
  pl.n = n;
  pl.nf = 0;
  int m = n;
  while (m % 4 == 0 && pl.nf < FFT_MAXF) { pl.fac[pl.nf++] = 4; m /= 4; }
  while (m % 2 == 0 && pl.nf < FFT_MAXF) { pl.fac[pl.nf++] = 2; m /= 2; }
  for (int f = 3; m > 1 && pl.nf < FFT_MAXF; f += 2)
    while (m % f == 0) { pl.fac[pl.nf++] = f; m /= f; }
}

void fft_twiddles(int n, double2* out) {
  for (int k = 0; k < n; ++k) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
    out[k] = make_double2((double)cosl(a), (double)sinl(a));
  }
}

hipEvent_t take_event(kb_ctx* c) {
  if (!c->freeEv.empty()) {
    hipEvent_t e = c->freeEv.back();
    c->freeEv.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

StageTimer::StageTimer(kb_ctx* c_, int s) : c(c_), stage(s) {
  if (c->timing) {
    ev.a = take_event(c);
    ev.b = take_event(c);
    if (ev.a) (void)hipEventRecord(ev.a, c->stream);
  }
}

StageTimer::~StageTimer() {
  if (c->timing && ev.a && ev.b) {
    (void)hipEventRecord(ev.b, c->stream);
    c->pending[stage].push_back(ev);
  }
}

int collect_timing(kb_ctx* c) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < NSTAGE; ++s) {
    for (auto& ev : c->pending[s]) {
      float ms = 0.f;
      HIPCHK(c, hipEventElapsedTime(&ms, ev.a, ev.b));
      c->tsum[s] += ms;
      c->tcount[s] += 1;
      c->freeEv.push_back(ev.a);
      c->freeEv.push_back(ev.b);
    }
    c->pending[s].clear();
  }
  return KB_OK;
}

// expandSimData! (SimDataStruct.jl:262-278) on the host mirror
void expand(kb_ctx* c) { advance_state(c->st, c->N); }

int upload_state(kb_ctx* c) {
  HIPCHK(c, hipMemcpyAsync(c->st_dev, &c->st, sizeof(StepState), hipMemcpyHostToDevice, c->stream));
  return KB_OK;
}

void drop_graph(kb_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
}

double hartree_prefactor(const kb_ctx* c) {
  if (c->p.pexp == KB_PEXP_NINVERSE_LO) return c->p.Lambda * (1.0 / 6.0);
  return c->p.Lambda * (double)(c->p.ONgroup + 2) / (6.0 * (double)c->p.ONgroup);
}

int launch_hartree(kb_ctx* c, int i, int j) {
  const double H = hartree_prefactor(c);
  const double hfac = H * (1.0 / (double)c->Vl);
  const long pr = pair_off(slot_of(c, i), slot_of(c, j), c->N);
  hipLaunchKernelGGL(hartree_kernel, dim3(1), dim3(1024), 0, c->stream, c->H, c->NP, c->V, pr, hfac, c->p.Mass2,
                     c->wgt_dev, c->m2);
  LAUNCHCHK(c);
  return KB_OK;
}

// the orbit products' arguments for the step in st
OrbitArgs orbit_args(kb_ctx* c, const StepState* st, double preF, double preR) {
  OrbitArgs o{};
  o.classical = c->p.pexp == KB_PEXP_LAMBDA_NLO_CLASSICAL;
  o.Mo = c->Mo;
  o.H = c->H;
  o.Y = c->Yo;
  o.G = c->G;
  o.S = c->S;
  o.st = st;
  o.NP = c->NP;
  o.N = c->N;
  o.V = c->V;
  o.Mp = c->Mp;
  o.Kp = c->Kp;
  o.ldg = c->ldg;
  o.preF = preF;
  o.preR = preR;
  o.X = c->Xo;
  o.PF = c->PF;
  o.PR = c->PR;
  o.Iring = c->Iring;
  o.PNP = c->PNP;
  o.ivol = 1.0 / (double)c->Vl;
  return o;
}

int launch_sigma(kb_ctx* c, double2* X, double* S_out) {
  const int rows = c->N - 1;  // grid rows; rows q >= L exit at once
  const double ivol = 1.0 / (double)c->Vl;
  const double lam = c->p.Lambda, ON = (double)c->p.ONgroup;
  SigmaArgs a{};
  a.H = c->H;
  a.S = S_out ? S_out : c->S;
  a.smap = c->smap_dev;
  a.rmap = c->rmap_dev;
  a.W = c->W;
  a.tw = c->tw;
  a.plan = c->plan;
  a.NP = c->NP;
  a.N = c->N;
  a.V = c->Vl;  // the transforms run on the whole lattice
  a.n = c->p.Nx;
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.P = c->fftP;
  a.npencils = c->Vl / c->p.Nx;
  a.st = c->st_dev;
  a.preF = ivol * ivol * ivol * (-(lam * lam) * (ON + 2.0) / (18.0 * ON * ON));
  a.preR = ivol * ivol * ivol * (-(lam * lam) * (ON + 2.0) / (6.0 * ON * ON));
  a.X = X;
  a.classical = c->p.pexp == KB_PEXP_LAMBDA_NLO_CLASSICAL;
  const int L = rows;
  if (c->fft_tile == KB_FFT_ORBIT_PRODUCT) {
    if (!c->Mo || X || S_out) return seterr(c, KB_ERR_STATE, "self-energy: the orbit product serves the forward step only");
    OrbitArgs o = orbit_args(c, c->st_dev, a.preF, a.preR);
    hipError_t e = launch_orbit(o, ORBIT_SUNSET, c->stream);
    if (e != hipSuccess) return seterr(c, KB_ERR_HIP, "orbit self-energy launch failed: %s", hipGetErrorString(e));
    return KB_OK;
  }
  if (plane_workspace_missing(c)) return seterr(c, KB_ERR_STATE, "self-energy: plane workspace not allocated");
  if (sigma_planes_supported(c->p.Nx, c->p.sdim)) {
    hipError_t e = launch_sigma_planes(a, L, c->stream, c->fft_tile);
    if (e != hipSuccess) return seterr(c, KB_ERR_HIP, "self-energy launch failed: %s", hipGetErrorString(e));
    return KB_OK;
  }
  if (X) {
    // the adjoint's recompute keeps the forward transforms: the pencil passes
    // that save them (2D and 3D past the plane kernels)
    hipError_t e = launch_sunset_pencils(a, L, c->stream);
    if (e != hipSuccess) return seterr(c, KB_ERR_HIP, "self-energy launch failed: %s", hipGetErrorString(e));
    return KB_OK;
  }
  const size_t lds = sizeof(double2) * 2 * (size_t)a.P * a.n;
  const dim3 grid((a.npencils + a.P - 1) / a.P, L);
  std::vector<std::pair<int, int>> passes;  // (axis, flags)
  const int sd = c->p.sdim;
  if (sd == 1) {
    passes.push_back({0, SIGMA_GATHER | SIGMA_SUNSET | SIGMA_SCATTER});
  } else if (sd == 2) {
    passes.push_back({0, SIGMA_GATHER});
    passes.push_back({1, SIGMA_SUNSET});
    passes.push_back({0, SIGMA_INVERSE | SIGMA_SCATTER});
  } else {
    passes.push_back({0, SIGMA_GATHER});
    passes.push_back({1, 0});
    passes.push_back({2, SIGMA_SUNSET});
    passes.push_back({1, SIGMA_INVERSE});
    passes.push_back({0, SIGMA_INVERSE | SIGMA_SCATTER});
  }
  for (auto& ps : passes) {
    a.axis = ps.first;
    a.flags = ps.second;
    hipLaunchKernelGGL(sigma_pass, grid, dim3(256), lds, c->stream, a);
    LAUNCHCHK(c);
  }
  return KB_OK;
}

// 1/N expansion: the bubble pass of row t-1 of the state `st` (writes row
// t-1 of the bubble triangles; X, if set, keeps the transforms), the
// resummation sweeps, and the chain pass that turns (IF, Ir) into Sigma
SigmaArgs ninv_sigma_args(kb_ctx* c, const StepState* st) {
  const double ivol = 1.0 / (double)c->Vl;
  SigmaArgs a{};
  a.H = c->H;
  a.S = c->S;
  a.smap = c->smap_dev;
  a.rmap = c->rmap_dev;
  a.W = c->W;
  a.tw = c->tw;
  a.plan = c->plan;
  a.NP = c->NP;
  a.N = c->N;
  a.V = c->Vl;
  a.n = c->p.Nx;
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.P = c->fftP;
  a.npencils = c->Vl / c->p.Nx;
  a.st = st;
  a.PF = c->PF;
  a.PR = c->PR;
  a.Iring = c->Iring;
  a.PNP = c->PNP;
  a.ivol = ivol;
  return a;
}

int launch_bubble(kb_ctx* c, const StepState* st, bool keep) {
  const double ivol = 1.0 / (double)c->Vl;
  if (c->fft_tile == KB_FFT_ORBIT_PRODUCT) {
    OrbitArgs o = orbit_args(c, st, ivol * ivol, ivol * ivol);
    if (!keep) o.X = nullptr;  // rows rebuilt outside a step keep no transforms
    hipError_t e = launch_orbit(o, ORBIT_BUBBLE, c->stream);
    if (e != hipSuccess) return seterr(c, KB_ERR_HIP, "orbit bubble launch failed: %s", hipGetErrorString(e));
    return KB_OK;
  }
  if (plane_workspace_missing(c)) return seterr(c, KB_ERR_STATE, "bubble pass: plane workspace not allocated");
  if (c->sym && !sigma_planes_supported(c->p.Nx, c->p.sdim))
    return seterr(c, KB_ERR_STATE, "bubble pass: orbit storage past the plane kernels runs the orbit products");
  SigmaArgs a = ninv_sigma_args(c, st);
  a.preF = a.preR = ivol * ivol;
  a.X = keep ? c->Xn : nullptr;
  hipError_t e = sigma_planes_supported(c->p.Nx, c->p.sdim) ? launch_bubble_planes(a, c->N - 1, c->stream)
                                                            : launch_bubble_pencils(a, c->N - 1, c->stream);
  if (e != hipSuccess) return seterr(c, KB_ERR_HIP, "bubble launch failed: %s", hipGetErrorString(e));
  return KB_OK;
}

int launch_resum(kb_ctx* c, const StepState* st, int passes) {
  NinvArgs a{};
  a.PF = c->PF;
  a.PR = c->PR;
  a.IR = c->IR;
  a.Iring = c->Iring;
  a.items = c->pitems;
  a.nitems = c->npitems;
  a.PNP = c->PNP;
  a.N = c->N;
  a.V = c->V;
  a.dt = c->p.dt;
  a.lam = c->p.Lambda;
  a.passes = passes;
  a.st = st;
  const size_t lds = ninv_lds(c->N);
  if (c->nsplit > 1) {
    const int ns = c->nsplit;
    const dim3 g(c->NB * ns), gf(c->NB);
    hipLaunchKernelGGL(ninv_split_sweep<1>, g, dim3(NINV_THREADS), lds, c->stream, a, ns, c->npart, c->ncv);
    hipLaunchKernelGGL(ninv_split_finish<1>, gf, dim3(NINV_FINISH_THREADS), lds, c->stream, a, ns, c->npart, c->ncv);
    LAUNCHCHK(c);
    if (passes == NINV_PASS_IR) return KB_OK;
    hipLaunchKernelGGL(ninv_split_sweep<2>, g, dim3(NINV_THREADS), lds, c->stream, a, ns, c->npart, c->ncv);
    hipLaunchKernelGGL(ninv_split_finish<2>, gf, dim3(NINV_FINISH_THREADS), lds, c->stream, a, ns, c->npart, c->ncv);
    hipLaunchKernelGGL(ninv_split_sweep<3>, g, dim3(NINV_THREADS), lds, c->stream, a, ns, c->npart, c->ncv);
    hipLaunchKernelGGL(ninv_split_finish<3>, gf, dim3(NINV_FINISH_THREADS), lds, c->stream, a, ns, c->npart, c->ncv);
    LAUNCHCHK(c);
    return KB_OK;
  }
  hipLaunchKernelGGL(ninv_resum_kernel, dim3(c->NB), dim3(NINV_THREADS), lds, c->stream, a);
  LAUNCHCHK(c);
  return KB_OK;
}

int launch_chain(kb_ctx* c) {
  const double ivol = 1.0 / (double)c->Vl;
  const double pre = -c->p.Lambda / (3.0 * (double)c->p.ONgroup) * ivol;
  if (c->fft_tile == KB_FFT_ORBIT_PRODUCT) {
    hipError_t e = launch_orbit(orbit_args(c, c->st_dev, pre, pre), ORBIT_CHAIN, c->stream);
    if (e != hipSuccess) return seterr(c, KB_ERR_HIP, "orbit chain launch failed: %s", hipGetErrorString(e));
    return KB_OK;
  }
  if (plane_workspace_missing(c)) return seterr(c, KB_ERR_STATE, "chain pass: plane workspace not allocated");
  if (c->sym && !sigma_planes_supported(c->p.Nx, c->p.sdim))
    return seterr(c, KB_ERR_STATE, "chain pass: orbit storage past the plane kernels runs the orbit products");
  SigmaArgs a = ninv_sigma_args(c, c->st_dev);
  a.preF = a.preR = pre;
  a.X = c->Xn;
  hipError_t e = sigma_planes_supported(c->p.Nx, c->p.sdim) ? launch_chain_planes(a, c->N - 1, c->stream)
                                                            : launch_chain_pencils(a, c->N - 1, c->stream);
  if (e != hipSuccess) return seterr(c, KB_ERR_HIP, "chain launch failed: %s", hipGetErrorString(e));
  return KB_OK;
}

// calcSelfEnergies! NinverseNLO (Evolution.jl:144-154) for the step in st_dev
int launch_ninv_stage(kb_ctx* c) {
  int rc;
  if ((rc = launch_bubble(c, c->st_dev, true))) return rc;
  {
    StageTimer st(c, 5);
    if ((rc = launch_resum(c, c->st_dev, NINV_PASS_ALL))) return rc;
  }
  return launch_chain(c);
}

// Rows tmin .. T-1 of the bubble and Ir triangles for the step the host
// window has just been advanced to (T = t - 1), one row after the other:
// needed after init and after any write to the window from outside a step.
int ninv_rebuild(kb_ctx* c) {
  const int T = c->st.t - 1;
  const int tmin = c->st.tmin;
  std::vector<StepState> fs;
  for (int R = tmin; R <= T - 1; ++R) {
    StepState f = c->st;
    f.t = R + 1;
    f.tmax = R + 1;
    f.L = R + 1 - tmin;
    f.s_t = R % c->N;
    f.s_tm1 = (R - 1) % c->N;
    f.s_tm2 = (R - 2 + c->N) % c->N;
    f.dead = -1;
    fs.push_back(f);
  }
  if (!fs.empty()) {
    HIPCHK(c, hipMemcpyAsync(c->stfill_dev, fs.data(), sizeof(StepState) * fs.size(), hipMemcpyHostToDevice,
                             c->stream));
    int rc;
    for (size_t k = 0; k < fs.size(); ++k) {
      if ((rc = launch_bubble(c, c->stfill_dev + k, false))) return rc;
      if ((rc = launch_resum(c, c->stfill_dev + k, NINV_PASS_IR))) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  c->ninv_valid = true;
  return KB_OK;
}

MemintArgs memint_args(const kb_ctx* c) {
  MemintArgs a{};
  a.H = c->H;
  a.S = c->S;
  a.kL2 = c->kL2;
  a.m2 = c->m2;
  a.NP = c->NP;
  a.N = c->N;
  a.V = c->V;
  a.dt = c->p.dt;
  a.integral = has_sigma(c);
  int ri = 1;
  for (int i = 0; i < 3; ++i)
    if (MEMINT_ROW_CHOICES[i] == kMemintVariants[c->tile].rows) ri = i;
  a.items = c->items_r[ri];
  a.nitems = c->nitems_r[ri];
  a.st = c->st_dev;
  return a;
}

hipError_t launch_memint(int tile, unsigned grid, size_t lds, hipStream_t s, const MemintArgs* args,
                         const int* wg_start, int nruns) {
  void* kargs[] = {(void*)&args, (void*)&wg_start, (void*)&nruns};
  return hipLaunchKernel(kMemintVariants[tile].fn, dim3(grid), dim3(kMemintVariants[tile].threads), kargs, lds, s);
}

size_t memint_lds(int N) { return sizeof(double) * 2 * (size_t)N * MB; }

int launch_hartree_step(kb_ctx* c) {
  const double H = hartree_prefactor(c);
  const double hfac = H * (1.0 / (double)c->Vl);
  hipLaunchKernelGGL(hartree_step_kernel, dim3(hartree_groups(c->V)), dim3(HARTREE_THREADS), 0, c->stream, c->H, c->NP,
                     c->V, c->N, c->st_dev, hfac, c->p.Mass2, c->wgt_dev, c->m2, c->hpart, nullptr);
  LAUNCHCHK(c);
  return KB_OK;
}

// Split memory integral (orbit storage): partial sweeps over item ranges,
// then the sums and the leapfrog, in the context's tile shape (the item
// table in margs_dev has that shape's row height).
int launch_memint_split(kb_ctx* c) {
  const MemintVariant& v = kMemintVariants[c->tile];
  const MemintArgs* margs = c->margs_dev;
  int ns = c->msplit;
  double* part = c->mpart;
  void* kargs[] = {(void*)&margs, (void*)&ns, (void*)&part};
  HIPCHK(c, hipLaunchKernel(v.partial, dim3(c->NB * c->msplit), dim3(v.threads), kargs, memint_lds(c->N), c->stream));
  hipLaunchKernelGGL(memint_finish, dim3(c->NB), dim3(MEMINT_FINISH_THREADS), memint_lds(c->N), c->stream,
                     c->margs_dev, c->msplit, c->mpart);
  LAUNCHCHK(c);
  return KB_OK;
}

// The kernels of one KB step (evolve!, Evolution.jl:924-981); every
// argument is step-invariant, the window comes from st_dev.
int enqueue_step(kb_ctx* c) {
  int rc;
  hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(64), 0, c->stream, c->stptr_dev, c->N_dev, 1);
  LAUNCHCHK(c);
  if (sunset_expansion(c->p.pexp)) {
    StageTimer st(c, 1);
    if ((rc = launch_sigma(c))) return rc;
  } else if (c->p.pexp == KB_PEXP_NINVERSE_NLO) {
    StageTimer st(c, 1);
    if ((rc = launch_ninv_stage(c))) return rc;
  }
  {
    StageTimer st(c, 2);
    if ((rc = launch_hartree_step(c))) return rc;
  }
  {
    StageTimer st(c, 0);
    if (c->msplit > 1) {
      if ((rc = launch_memint_split(c))) return rc;
    } else {
      hipError_t e = launch_memint(c->tile, c->NB, memint_lds(c->N), c->stream, c->margs_dev, c->wg0_dev, 1);
      if (e != hipSuccess) return seterr(c, KB_ERR_HIP, "memory-integral launch failed: %s", hipGetErrorString(e));
    }
  }
  return KB_OK;
}

int upload(kb_ctx* c, double* dst, const double* src, size_t n) {
  HIPCHK(c, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return KB_OK;
}

int download(kb_ctx* c, double* dst, const double* src, size_t n) {
  HIPCHK(c, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return KB_OK;
}

int lat_upload(kb_ctx* c, double* dst, const double* src) {
  if (!c->sym) return upload(c, dst, src, c->V);
  c->lat_tmp.resize(c->V);
  for (int r = 0; r < c->V; ++r) c->lat_tmp[r] = src[c->rep_site[r]];
  HIPCHK(c, hipMemcpyAsync(dst, c->lat_tmp.data(), sizeof(double) * c->V, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));  // lat_tmp is reused by the next call
  return KB_OK;
}

int lat_download(kb_ctx* c, double* dst, const double* src) {
  if (!c->sym) return download(c, dst, src, c->V);
  c->lat_tmp.resize(c->V);
  int rc = download(c, c->lat_tmp.data(), src, c->V);
  if (rc) return rc;
  for (int s = 0; s < c->Vl; ++s) dst[s] = c->lat_tmp[c->orbit_of[s]];
  return KB_OK;
}

void drop_graph(kb_ctx* c);

void free_all(kb_ctx* c) {
  drop_graph(c);
  for (void* ptr : {(void*)c->H, (void*)c->S, (void*)c->W, (void*)c->kL2, (void*)c->m2, (void*)c->hpart, (void*)c->partial,
                    (void*)c->modes, (void*)c->vbuf, (void*)c->tw, (void*)c->flag, (void*)c->items_r[0],
                    (void*)c->items_r[1], (void*)c->items_r[2], (void*)c->smap_dev, (void*)c->rmap_dev,
                    (void*)c->wgt_dev, (void*)c->PF, (void*)c->PR, (void*)c->IR, (void*)c->Iring, (void*)c->Xn,
                    (void*)c->stfill_dev, (void*)c->pitems, (void*)c->mpart, (void*)c->npart, (void*)c->ncv, (void*)c->Mo, (void*)c->Yo, (void*)c->Xo, (void*)c->G,
                    (void*)c->st_dev, (void*)c->stptr_dev, (void*)c->N_dev, (void*)c->margs_dev, (void*)c->wg0_dev})
    if (ptr) (void)hipFree(ptr);
  adj_free(c);
  for (int s = 0; s < NSTAGE; ++s)
    for (auto& ev : c->pending[s]) {
      (void)hipEventDestroy(ev.a);
      (void)hipEventDestroy(ev.b);
    }
  for (auto e : c->freeEv) (void)hipEventDestroy(e);
  if (c->own) (void)hipStreamDestroy(c->own);
}

int dmalloc(kb_ctx* c, void** ptr, size_t bytes) {
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 8);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return seterr(c, KB_ERR_OOM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  }
  c->bytes += bytes;
  return KB_OK;
}

int meas_grid(const kb_ctx* c) { return (c->V + MEAS_THREADS - 1) / MEAS_THREADS; }

// The plane kernels' workspace: W holds one complex lattice per row between
// the passes (3D lattices, and the pencil passes past Nx = 64); Xn keeps the
// 1/N bubble pass's transforms for its chain pass.  Allocated when a plane
// form is first selected: LambdaLO and contexts on the orbit products never
// hold them.
int ensure_plane_workspace(kb_ctx* c) {
  if (!has_sigma(c)) return KB_OK;
  int rc;
  const size_t rows = (size_t)c->N * c->Vl * sizeof(double2);
  // the plane passes keep the half space z in [0, Nx/2] of each row; the
  // pencil passes (Nx > 64) whole rows
  const int n = c->p.Nx;
  const size_t wrows = sigma_planes_supported(n, c->p.sdim) ? (size_t)c->N * n * n * (n / 2 + 1) * sizeof(double2) : rows;
  if ((c->p.sdim == 3 || !sigma_planes_supported(n, c->p.sdim)) && !c->W && (rc = dmalloc(c, (void**)&c->W, wrows)))
    return rc;
  if (c->p.pexp == KB_PEXP_NINVERSE_NLO && !c->Xn && (rc = dmalloc(c, (void**)&c->Xn, rows))) return rc;
  return KB_OK;
}

// a plane-kernel stage needs its workspace (a guard, not a path callers reach)
bool plane_workspace_missing(const kb_ctx* c) {
  return ((c->p.sdim == 3 || !sigma_planes_supported(c->p.Nx, c->p.sdim)) && !c->W) ||
         (c->p.pexp == KB_PEXP_NINVERSE_NLO && !c->Xn);
}

void cubic_orbits(int Nx, int sdim, std::vector<int>& orbit_of, std::vector<int>& rep_site) {
  int Vl = 1;
  for (int d = 0; d < sdim; ++d) Vl *= Nx;
  orbit_of.assign(Vl, -1);
  rep_site.clear();
  std::vector<int> key_to_orbit;  // canonical key (sorted folded coordinates) -> orbit
  const int half = Nx / 2 + 1;
  int nkeys = 1;
  for (int d = 0; d < sdim; ++d) nkeys *= half;
  key_to_orbit.assign(nkeys, -1);
  for (int s = 0; s < Vl; ++s) {
    int m[3] = {0, 0, 0}, rem = s;
    for (int d = 0; d < sdim; ++d) {
      const int n = rem % Nx;
      rem /= Nx;
      m[d] = n < Nx - n ? n : Nx - n;  // reflection
    }
    for (int i = 0; i < sdim; ++i)  // permutation: sort
      for (int j = i + 1; j < sdim; ++j)
        if (m[j] < m[i]) {
          const int x = m[i];
          m[i] = m[j];
          m[j] = x;
        }
    int key = 0;
    for (int d = sdim - 1; d >= 0; --d) key = key * half + m[d];
    if (key_to_orbit[key] < 0) {
      key_to_orbit[key] = (int)rep_site.size();
      rep_site.push_back(s);
    }
    orbit_of[s] = key_to_orbit[key];
  }
}

}  // namespace
#pragma GCC visibility pop

extern "C" {

const char* kb_last_error(const kb_ctx* ctx) { return ctx ? ctx->err.c_str() : free_call_error(); }

namespace {

// Orbits of the cubic group of the lattice (axis permutations and
// reflections n_i -> Nx - n_i) on the sites, numbered in order of first
// appearance; the first site of an orbit is its representative.  kL2 and
// the whole KB step are invariant under the group, so a cubic-symmetric
// state is carried exactly by one value per orbit.
static int create_ctx(const kb_params* p, bool sym, kb_ctx** out);

}  // namespace

int kb_create(const kb_params* p, kb_ctx** out) { return create_ctx(p, false, out); }

int kb_create_sym(const kb_params* p, kb_ctx** out) { return create_ctx(p, true, out); }

int kb_storage_info(const kb_ctx* c, long long* stored_sites, long long* lattice_sites) {
  if (!c) return KB_ERR_ARG;
  if (stored_sites) *stored_sites = c->V;
  if (lattice_sites) *lattice_sites = c->Vl;
  return KB_OK;
}

}  // extern "C"

namespace {

static int create_ctx(const kb_params* p, bool sym, kb_ctx** out) {
  if (!p || !out) return KB_ERR_ARG;
  *out = nullptr;
  if (p->sdim < 1 || p->sdim > 3 || p->Nx < 1 || p->Nx > KB_MAX_NX || p->NstepsinMemory < 3 ||
      p->NstepsinMemory > 640 || p->ONgroup < 1 || !(p->dt > 0) || p->pexp < KB_PEXP_LAMBDA_LO ||
      p->pexp > KB_PEXP_LAMBDA_NLO_CLASSICAL)
    return KB_ERR_ARG;
  {
    long v = 1;
    for (int d = 0; d < p->sdim; ++d) v *= p->Nx;
    if (v > KB_MAX_SITES) return KB_ERR_ARG;
  }
  kb_ctx* c = new (std::nothrow) kb_ctx();
  if (!c) return KB_ERR_OOM;
  c->p = *p;
  c->N = p->NstepsinMemory;
  c->NP = slot_pairs(c->N);
  c->Vl = 1;
  for (int d = 0; d < p->sdim; ++d) c->Vl *= p->Nx;
  c->sym = sym;
  if (sym) {
    cubic_orbits(p->Nx, p->sdim, c->orbit_of, c->rep_site);
    c->V = (int)c->rep_site.size();
  } else {
    c->V = c->Vl;
  }
  c->nx = p->Nx;
  c->ny = p->sdim >= 2 ? p->Nx : 1;
  c->nz = p->sdim >= 3 ? p->Nx : 1;
  c->NB = (c->V + MB - 1) / MB;
  c->Vpad = c->NB * MB;
  plan_fft(c->plan, p->Nx);
  // pencils per workgroup: 16 complex elements per pencil row at least, and
  // the two LDS buffers (2 P Nx complex values) within 64 KiB
  c->fftP = p->Nx >= 64 ? 4 : (p->Nx >= 16 ? 16 : 32);
  c->fftP = std::min(c->fftP, std::max(1, KB_MAX_NX / p->Nx));
  if (c->fftP > c->Vl / p->Nx) c->fftP = c->Vl / p->Nx;
  int rc;
  auto fail = [&](int code) {
    free_all(c);
    delete c;
    return code;
  };
  // orbit storage gathers and scatters through the site maps in the plane
  // kernels only (Nx <= 64); past them, a transform stage needs the orbit
  // products (LambdaLO has none)
  if (sym && !sigma_planes_supported(p->Nx, p->sdim) && (sunset_expansion(p->pexp) || p->pexp == KB_PEXP_NINVERSE_NLO) &&
      c->V > KB_ORBIT_PRODUCT_MAX)
    return fail(KB_ERR_ARG);
  if (hipSetDevice(p->device) != hipSuccess) return fail(KB_ERR_HIP);
  if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) return fail(KB_ERR_HIP);
  c->stream = c->own;
  const size_t hcount = (size_t)c->NB * (size_t)c->NP * MB * 2;
  if ((rc = dmalloc(c, (void**)&c->H, hcount * sizeof(double)))) return fail(rc);
  if ((rc = dmalloc(c, (void**)&c->S, (size_t)c->NB * c->N * MB * 2 * sizeof(double)))) return fail(rc);
  c->PNP = pslot_pairs(c->N);
  const size_t pcount = (size_t)c->NB * (size_t)c->PNP * MB;  // one value per pair and site
  const size_t rcount = (size_t)c->NB * c->N * MB * 2;
  if (p->pexp == KB_PEXP_NINVERSE_NLO) {
    if ((rc = dmalloc(c, (void**)&c->PF, pcount * sizeof(double))) ||
        (rc = dmalloc(c, (void**)&c->PR, pcount * sizeof(double))) ||
        (rc = dmalloc(c, (void**)&c->IR, pcount * sizeof(double))) ||
        (rc = dmalloc(c, (void**)&c->Iring, rcount * sizeof(double))) ||
        (rc = dmalloc(c, (void**)&c->stfill_dev, (size_t)c->N * sizeof(StepState))))
      return fail(rc);
    std::vector<int2> items;
    for (int J = 0; J * PCB < c->N; ++J)
      for (int k = 0; J * PCB + k * NINV_ROWS < c->N; ++k) items.push_back(make_int2(J, k));
    c->npitems = (int)items.size();
    if ((rc = dmalloc(c, (void**)&c->pitems, sizeof(int2) * items.size()))) return fail(rc);
    if (hipMemcpy(c->pitems, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice) != hipSuccess)
      return fail(KB_ERR_HIP);
  }
  if (sym) {
    std::vector<int> rmap(c->Vl, -1);
    std::vector<double> wgt(c->V, 0.0);
    for (int s = 0; s < c->Vl; ++s) wgt[c->orbit_of[s]] += 1.0;
    for (int r = 0; r < c->V; ++r) rmap[c->rep_site[r]] = r;
    if ((rc = dmalloc(c, (void**)&c->smap_dev, sizeof(int) * c->Vl)) ||
        (rc = dmalloc(c, (void**)&c->rmap_dev, sizeof(int) * c->Vl)) ||
        (rc = dmalloc(c, (void**)&c->wgt_dev, sizeof(double) * c->V)))
      return fail(rc);
    if (hipMemcpy(c->smap_dev, c->orbit_of.data(), sizeof(int) * c->Vl, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->rmap_dev, rmap.data(), sizeof(int) * c->Vl, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->wgt_dev, wgt.data(), sizeof(double) * c->V, hipMemcpyHostToDevice) != hipSuccess)
      return fail(KB_ERR_HIP);
  }
  // orbit storage, LambdaNLO: the self-energy stage as products with the
  // orbit cosine matrix Mo[o][b] = sum_{x in orbit b} cos(2 pi k_o . x / Nx),
  // k_o the representative of orbit o (zero-padded to the product's tiles)
  if (sym && (sunset_expansion(p->pexp) || p->pexp == KB_PEXP_NINVERSE_NLO) && c->V <= KB_ORBIT_PRODUCT_MAX) {
    c->Mp = (c->V + ORBIT_TILE - 1) / ORBIT_TILE * ORBIT_TILE;
    c->Kp = (c->V + ORBIT_K - 1) / ORBIT_K * ORBIT_K;
    std::vector<double> mo((size_t)c->Mp * c->Kp, 0.0), ctab(p->Nx);
    for (int j = 0; j < p->Nx; ++j)
      ctab[j] = (double)cosl(2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)p->Nx);
    std::vector<int> xc((size_t)c->Vl * p->sdim);
    for (int x = 0; x < c->Vl; ++x)
      for (int d = 0, rem = x; d < p->sdim; ++d, rem /= p->Nx) xc[(size_t)x * p->sdim + d] = rem % p->Nx;
    for (int o = 0; o < c->V; ++o) {
      const int* ko = &xc[(size_t)c->rep_site[o] * p->sdim];
      double* row = &mo[(size_t)o * c->Kp];
      for (int x = 0; x < c->Vl; ++x) {
        int m = 0;
        for (int d = 0; d < p->sdim; ++d) m += ko[d] * xc[(size_t)x * p->sdim + d];
        row[c->orbit_of[x]] += ctab[m % p->Nx];
      }
    }
    // Y and G: Kp rows x whole column tiles, zero outside the live block
    c->ldg = (2 * (c->N - 1) + ORBIT_TILE - 1) / ORBIT_TILE * ORBIT_TILE;
    const size_t ybytes = (size_t)c->Kp * c->ldg * sizeof(double);
    if ((rc = dmalloc(c, (void**)&c->Mo, mo.size() * sizeof(double))) ||
        (rc = dmalloc(c, (void**)&c->Yo, ybytes)) || (rc = dmalloc(c, (void**)&c->G, ybytes)))
      return fail(rc);
    if (hipMemset(c->Yo, 0, ybytes) != hipSuccess || hipMemset(c->G, 0, ybytes) != hipSuccess)
      return fail(KB_ERR_HIP);
    if (p->pexp == KB_PEXP_NINVERSE_NLO) {
      if ((rc = dmalloc(c, (void**)&c->Xo, ybytes))) return fail(rc);
      if (hipMemset(c->Xo, 0, ybytes) != hipSuccess) return fail(KB_ERR_HIP);
    }
    if (hipMemcpy(c->Mo, mo.data(), mo.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return fail(KB_ERR_HIP);
    if (orbit_prepare() != hipSuccess) {
      seterr(c, KB_ERR_ARG, "orbit product kernels: LDS attribute rejected");
      return fail(KB_ERR_ARG);
    }
    c->fft_tile = KB_FFT_ORBIT_PRODUCT;
  }
  if (c->fft_tile != KB_FFT_ORBIT_PRODUCT && (rc = ensure_plane_workspace(c))) return fail(rc);
  if ((rc = dmalloc(c, (void**)&c->kL2, (size_t)c->Vpad * sizeof(double)))) return fail(rc);
  if ((rc = dmalloc(c, (void**)&c->m2, 8 * sizeof(double)))) return fail(rc);
  if ((rc = dmalloc(c, (void**)&c->hpart, hartree_part_doubles(c->V) * sizeof(double)))) return fail(rc);
  if ((rc = dmalloc(c, (void**)&c->partial, (size_t)6 * meas_grid(c) * sizeof(double)))) return fail(rc);
  if ((rc = dmalloc(c, (void**)&c->modes, (size_t)6 * c->V * sizeof(double)))) return fail(rc);
  // staging: 4 lattices, or the 4 site-1 history vectors of kb_measure_history
  const size_t vlen = (size_t)(c->V > c->N ? c->V : c->N);
  if ((rc = dmalloc(c, (void**)&c->vbuf, 4 * vlen * sizeof(double)))) return fail(rc);
  if ((rc = dmalloc(c, (void**)&c->tw, (size_t)p->Nx * sizeof(double2)))) return fail(rc);
  if ((rc = dmalloc(c, (void**)&c->flag, sizeof(int) * 4))) return fail(rc);
  // memory-integral work items: chunks of MEMINT_ROWS rows of each column
  // block, in storage order
  for (int i = 0; i < 3; ++i) {
    const int rows = MEMINT_ROW_CHOICES[i];
    std::vector<int2> items;
    for (int J = 0; J * CB < c->N; ++J)
      for (int k = 0; J * CB + k * rows < c->N; ++k) items.push_back(make_int2(J, k));
    c->nitems_r[i] = (int)items.size();
    if ((rc = dmalloc(c, (void**)&c->items_r[i], sizeof(int2) * items.size()))) return fail(rc);
    if (hipMemcpy(c->items_r[i], items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice) != hipSuccess)
      return fail(KB_ERR_HIP);
    if (rows == MEMINT_ROWS) {
      c->items = c->items_r[i];
      c->nitems = c->nitems_r[i];
    }
  }
  if ((rc = dmalloc(c, (void**)&c->st_dev, sizeof(StepState)))) return fail(rc);
  if ((rc = dmalloc(c, (void**)&c->stptr_dev, sizeof(StepState*)))) return fail(rc);
  if ((rc = dmalloc(c, (void**)&c->N_dev, sizeof(int)))) return fail(rc);
  if ((rc = dmalloc(c, (void**)&c->margs_dev, sizeof(MemintArgs)))) return fail(rc);
  if ((rc = dmalloc(c, (void**)&c->wg0_dev, sizeof(int)))) return fail(rc);
  {
    const int zero = 0;
    const MemintArgs ma = memint_args(c);
    if (hipMemcpy(c->st_dev, &c->st, sizeof(StepState), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->stptr_dev, &c->st_dev, sizeof(StepState*), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->N_dev, &c->N, sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->margs_dev, &ma, sizeof(MemintArgs), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->wg0_dev, &zero, sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      return fail(KB_ERR_HIP);
  }
  if (sigma_planes_supported(p->Nx, p->sdim) && sigma_planes_prepare(p->Nx) != hipSuccess) {
    seterr(c, KB_ERR_ARG, "self-energy plane kernels: LDS attribute rejected");
    return fail(KB_ERR_ARG);
  }
  {
    // the attribute is per function, shared by every context: allow the
    // largest supported window, not this context's
    const size_t lds = memint_lds(640);
    hipError_t e = hipSuccess;
    for (int v = 0; v < NMEMINT_VARIANTS && e == hipSuccess; ++v) {
      e = hipFuncSetAttribute(kMemintVariants[v].fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e == hipSuccess)
        e = hipFuncSetAttribute(kMemintVariants[v].partial, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    }
    if (e == hipSuccess)
      e = hipFuncSetAttribute((const void*)memint_finish, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) {
      seterr(c, KB_ERR_ARG, "memory-integral kernel: %zu bytes of LDS rejected: %s", lds, hipGetErrorString(e));
      return fail(KB_ERR_ARG);
    }
  }
  // orbit storage leaves fewer site blocks than workgroup slots: split each
  // block's memory integral over msplit workgroups, as many as fit the
  // device's slots for this window's LDS size in one round (a grid that
  // overfills some CUs by one workgroup waits for those CUs)
  if (sym) {
    int per_cu = 0;
    hipDeviceProp_t prop;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)memint_partial<64, 2, 256>, 256,
                                                     memint_lds(c->N)) != hipSuccess ||
        hipGetDeviceProperties(&prop, p->device) != hipSuccess)
      return fail(KB_ERR_HIP);
    const int slots = per_cu * prop.multiProcessorCount;
    c->msplit = std::max(1, std::min(KB_MSPLIT_MAX, slots / c->NB));
    if (c->msplit > 1 &&
        (rc = dmalloc(c, (void**)&c->mpart, (size_t)c->NB * c->msplit * 2 * c->N * MB * sizeof(double))))
      return fail(rc);
  }
  if (hipMemsetAsync(c->H, 0, hcount * sizeof(double), c->stream) != hipSuccess) return fail(KB_ERR_HIP);
  if (hipMemsetAsync(c->S, 0, (size_t)c->NB * c->N * MB * 2 * sizeof(double), c->stream) != hipSuccess)
    return fail(KB_ERR_HIP);
  if (hipMemsetAsync(c->m2, 0, 8 * sizeof(double), c->stream) != hipSuccess) return fail(KB_ERR_HIP);
  if (hipMemsetAsync(c->hpart, 0, hartree_part_doubles(c->V) * sizeof(double), c->stream) != hipSuccess)
    return fail(KB_ERR_HIP);
  if (c->PF) {
    for (double* ptr : {c->PF, c->PR, c->IR})
      if (hipMemsetAsync(ptr, 0, pcount * sizeof(double), c->stream) != hipSuccess) return fail(KB_ERR_HIP);
    if (hipMemsetAsync(c->Iring, 0, rcount * sizeof(double), c->stream) != hipSuccess) return fail(KB_ERR_HIP);
    const void* rfns[] = {(const void*)ninv_resum_kernel,      (const void*)ninv_split_sweep<1>,
                          (const void*)ninv_split_sweep<2>,    (const void*)ninv_split_sweep<3>,
                          (const void*)ninv_split_finish<1>,   (const void*)ninv_split_finish<2>,
                          (const void*)ninv_split_finish<3>};
    for (const void* f : rfns)
      if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ninv_lds(640)) != hipSuccess) {
        seterr(c, KB_ERR_ARG, "resummation kernels: LDS attribute rejected");
        return fail(KB_ERR_ARG);
      }
    // orbit storage: split the sweeps like the memory integral
    if (sym) {
      int per_cu = 0;
      hipDeviceProp_t prop;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)ninv_split_sweep<1>, NINV_THREADS,
                                                       ninv_lds(c->N)) != hipSuccess ||
          hipGetDeviceProperties(&prop, p->device) != hipSuccess)
        return fail(KB_ERR_HIP);
      c->nsplit = std::max(1, std::min(KB_MSPLIT_MAX, per_cu * prop.multiProcessorCount / c->NB));
      if (c->nsplit > 1 &&
          ((rc = dmalloc(c, (void**)&c->npart, (size_t)c->NB * c->nsplit * c->N * MB * sizeof(double))) ||
           (rc = dmalloc(c, (void**)&c->ncv, (size_t)c->NB * c->N * MB * sizeof(double)))))
        return fail(rc);
    }
  }
  // twiddles exp(-2 pi i k / Nx) and the default k^2 table
  std::vector<double2> tw(p->Nx);
  fft_twiddles(p->Nx, tw.data());
  std::vector<double> k2(c->Vpad, 0.0);
  for (int i = 0; i < c->V; ++i) {
    int rem = sym ? c->rep_site[i] : i;
    double s2 = 0.0;
    for (int d = 0; d < p->sdim; ++d) {
      const int nd = rem % p->Nx;
      rem /= p->Nx;
      const double s = std::sin(M_PI * (double)nd / (double)p->Nx);
      s2 += 4.0 * s * s;
    }
    k2[i] = s2;
  }
  if (hipMemcpy(c->tw, tw.data(), sizeof(double2) * p->Nx, hipMemcpyHostToDevice) != hipSuccess)
    return fail(KB_ERR_HIP);
  if (hipMemcpy(c->kL2, k2.data(), sizeof(double) * c->Vpad, hipMemcpyHostToDevice) != hipSuccess)
    return fail(KB_ERR_HIP);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(KB_ERR_HIP);
  *out = c;
  return KB_OK;
}

}  // namespace

extern "C" {

int kb_destroy(kb_ctx* ctx) {
  if (!ctx) return KB_OK;
  (void)hipSetDevice(ctx->p.device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_all(ctx);
  delete ctx;
  return KB_OK;
}

int kb_info(const kb_ctx* ctx, long long* V, long long* device_bytes) {
  if (!ctx) return KB_ERR_ARG;
  if (V) *V = ctx->Vl;
  if (device_bytes) *device_bytes = (long long)ctx->bytes;
  return KB_OK;
}

int kb_get_kL2(kb_ctx* c, double* out) {
  if (!c || !out) return KB_ERR_ARG;
  return lat_download(c, out, c->kL2);
}

int kb_set_kL2(kb_ctx* c, const double* in) {
  if (!c || !in) return KB_ERR_ARG;
  int rc = lat_upload(c, c->kL2, in);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return KB_OK;
}

int kb_init_rows(kb_ctx* c, const double* F11, const double* F21, const double* F22) {
  if (!c || !F11 || !F21 || !F22) return KB_ERR_ARG;
  if (c->st.tmax != 0) return seterr(c, KB_ERR_STATE, "init: context already initialised");
  int rc;
  double* d = c->vbuf;
  if ((rc = lat_upload(c, d, F11)) || (rc = lat_upload(c, d + c->V, F21)) || (rc = lat_upload(c, d + 2 * c->V, F22)))
    return rc;
  expand(c);
  expand(c);
  if ((rc = upload_state(c))) return rc;
  const int blocks = (c->V + 255) / 256;
  hipLaunchKernelGGL(init_rows_kernel, dim3(blocks), dim3(256), 0, c->stream, c->H, c->NP, c->V, c->N, d,
                     d + c->V, d + 2 * c->V, c->p.dt);
  LAUNCHCHK(c);
  if ((rc = launch_hartree(c, 2, 2))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->initialised = true;
  c->ninv_valid = false;
  return KB_OK;
}

int kb_init_particles(kb_ctx* c, const double* n, double initmass) {
  if (!c || !n) return KB_ERR_ARG;
  if (c->st.tmax != 0) return seterr(c, KB_ERR_STATE, "init: context already initialised");
  int rc;
  double* d = c->vbuf;
  if ((rc = lat_upload(c, d + 3 * c->V, n))) return rc;
  const int blocks = (c->V + 255) / 256;
  hipLaunchKernelGGL(init_particles_kernel, dim3(blocks), dim3(256), 0, c->stream, c->kL2, d + 3 * c->V,
                     initmass, c->p.dt, c->V, d, d + c->V, d + 2 * c->V);
  LAUNCHCHK(c);
  expand(c);
  expand(c);
  if ((rc = upload_state(c))) return rc;
  hipLaunchKernelGGL(init_rows_kernel, dim3(blocks), dim3(256), 0, c->stream, c->H, c->NP, c->V, c->N, d,
                     d + c->V, d + 2 * c->V, c->p.dt);
  LAUNCHCHK(c);
  if ((rc = launch_hartree(c, 2, 2))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->initialised = true;
  c->ninv_valid = false;
  c->initmass = initmass;
  c->init_particles = true;
  return KB_OK;
}

int kb_step(kb_ctx* c, int t) {
  if (!c) return KB_ERR_ARG;
  if (!c->initialised) return seterr(c, KB_ERR_STATE, "step before init");
  if (c->in_ensemble) return seterr(c, KB_ERR_STATE, "context is an ensemble member: use kb_ens_step");
  if (t != c->st.tmax + 1 || t < 3)
    return seterr(c, KB_ERR_WINDOW, "evolve: t=%d but window is [%d,%d]", t, c->st.tmin, c->st.tmax);
  expand(c);  // the device state is advanced by the step's first kernel
  if (c->p.pexp == KB_PEXP_NINVERSE_NLO && !c->ninv_valid) {
    int rc = ninv_rebuild(c);
    if (rc) return rc;
  }
  if (c->adj) {
    // recording for the adjoint: eager launches with the tape kept
    int rc;
    if ((rc = adj_before_step(c))) return rc;
    if ((rc = enqueue_step(c))) return rc;
    return adj_after_step(c);
  }
  if (c->use_graph && !c->timing) {
    if (!c->gexec) {
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int rc = enqueue_step(c);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return seterr(c, KB_ERR_HIP, "step capture failed: %s", hipGetErrorString(e));
      c->graph = g;
      HIPCHK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    }
    HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    return KB_OK;
  }
  return enqueue_step(c);
}

int kb_use_graph(kb_ctx* c, int on) {
  if (!c) return KB_ERR_ARG;
  if (!on) drop_graph(c);
  c->use_graph = on != 0;
  return KB_OK;
}

int kb_window(const kb_ctx* c, int* tmin, int* tmax) {
  if (!c) return KB_ERR_ARG;
  if (tmin) *tmin = c->st.tmin;
  if (tmax) *tmax = c->st.tmax;
  return KB_OK;
}

int kb_hartree(kb_ctx* c, double* m2) {
  if (!c || !m2) return KB_ERR_ARG;
  return download(c, m2, c->m2, 1);
}

int kb_measure(kb_ctx* c, int t, kb_meas* out, double* F, double* ddF, double* r, double* n, double* omega,
               double* cc) {
  if (!c || !out) return KB_ERR_ARG;
  if (!c->initialised) return seterr(c, KB_ERR_STATE, "measure before init");
  if (t != c->st.tmax || !inwin(c, t - 1)) return seterr(c, KB_ERR_WINDOW, "measure: t=%d, window [%d,%d]", t, c->st.tmin, c->st.tmax);
  MeasArgs a{};
  a.wgt = c->wgt_dev;
  a.H = c->H;
  a.S = c->S;
  a.kL2 = c->kL2;
  a.NP = c->NP;
  a.N = c->N;
  a.V = c->V;
  a.s_t = slot_of(c, t);
  a.s_tm1 = slot_of(c, t - 1);
  a.tmin = c->st.tmin;
  a.t = t;
  a.dt = c->p.dt;
  a.integral = has_sigma(c);
  const bool want = F || ddF || r || n || omega || cc;
  if (want) {
    a.outF = c->modes;
    a.outddF = c->modes + c->V;
    a.outr = c->modes + 2 * c->V;
    a.outn = c->modes + 3 * c->V;
    a.outom = c->modes + 4 * c->V;
    a.outc = c->modes + 5 * c->V;
  }
  a.partial = c->partial;
  const int g = meas_grid(c);
  hipLaunchKernelGGL(measure_kernel, dim3(g), dim3(MEAS_THREADS), 0, c->stream, a);
  LAUNCHCHK(c);
  std::vector<double> part((size_t)6 * g);
  int rc;
  if ((rc = download(c, part.data(), c->partial, part.size()))) return rc;
  double m2 = 0;
  if ((rc = download(c, &m2, c->m2, 1))) return rc;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int b = 0; b < g; ++b)
    for (int k = 0; k < 6; ++k) s[k] += part[(size_t)b * 6 + k];
  const double dt = c->p.dt;
  const double H = hartree_prefactor(c);
  const double ivol = 1.0 / (double)c->Vl;
  out->time = (double)(t - 1) * dt * c->p.Mass;
  out->E_kin = s[0];
  // E_pot = sum_p 1/2 (Mass2 + kL2 + 1/2 ivol H sum_q F) F  (Measurement.jl:135-149)
  out->E_pot = 0.5 * (c->p.Mass2 + 0.5 * ivol * H * s[1]) * s[1] + 0.5 * s[2];
  double Emem = 0.0;
  if (t != 2 && a.integral) {
    Emem = 0.5 * 0.25 * dt * s[3];
    if (t > 3) Emem += 0.25 * dt * s[4];
  }
  out->E_mem = Emem;
  out->E = out->E_kin + out->E_pot + out->E_mem;
  out->M2 = m2;
  out->n_unstable = (int)s[5];
  double* dsts[6] = {F, ddF, r, n, omega, cc};
  for (int k = 0; k < 6; ++k)
    if (dsts[k] && (rc = lat_download(c, dsts[k], c->modes + (size_t)k * c->V))) return rc;
  return KB_OK;
}

int kb_measure_history(kb_ctx* c, int t, double* SigF, double* Sigr, double* Fonet, double* ronet) {
  if (!c) return KB_ERR_ARG;
  if (t != c->st.tmax) return seterr(c, KB_ERR_WINDOW, "history: t=%d, window [%d,%d]", t, c->st.tmin, c->st.tmax);
  const int L = t - c->st.tmin;
  if (L <= 0) return KB_OK;
  hipLaunchKernelGGL(history_kernel, dim3((L + 255) / 256), dim3(256), 0, c->stream, c->H, c->S, c->NP, c->N,
                     c->st.tmin, t, c->vbuf);
  LAUNCHCHK(c);
  std::vector<double> h((size_t)4 * L);
  int rc;
  if ((rc = download(c, h.data(), c->vbuf, h.size()))) return rc;
  double* dsts[4] = {SigF, Sigr, Fonet, ronet};
  for (int k = 0; k < 4; ++k)
    if (dsts[k]) memcpy(dsts[k], h.data() + (size_t)k * L, sizeof(double) * L);
  return KB_OK;
}

int kb_get_entry(kb_ctx* c, int which, int i, int j, double* out) {
  if (!c || !out || (which != 0 && which != 1)) return KB_ERR_ARG;
  if (!inwin(c, i) || !inwin(c, j))
    return seterr(c, KB_ERR_WINDOW, "get: (%d,%d) outside window [%d,%d]", i, j, c->st.tmin, c->st.tmax);
  const int si = slot_of(c, i), sj = slot_of(c, j);
  const double sg = which ? (si >= sj ? 1.0 : -1.0) : 1.0;
  hipLaunchKernelGGL(gather_pair, dim3((c->V + 255) / 256), dim3(256), 0, c->stream, c->H, c->NP, c->V,
                     pair_off(si, sj, c->N), which, sg, c->vbuf);
  LAUNCHCHK(c);
  return lat_download(c, out, c->vbuf);
}

int kb_set_entry(kb_ctx* c, int which, int i, int j, const double* in) {
  if (!c || !in || (which != 0 && which != 1)) return KB_ERR_ARG;
  if (!inwin(c, i) || !inwin(c, j))
    return seterr(c, KB_ERR_WINDOW, "set: (%d,%d) outside window [%d,%d]", i, j, c->st.tmin, c->st.tmax);
  const int si = slot_of(c, i), sj = slot_of(c, j);
  const double sg = which ? (si >= sj ? 1.0 : -1.0) : 1.0;
  int rc;
  if ((rc = lat_upload(c, c->vbuf, in))) return rc;
  hipLaunchKernelGGL(scatter_pair, dim3((c->V + 255) / 256), dim3(256), 0, c->stream, c->H, c->NP, c->V,
                     pair_off(si, sj, c->N), which, sg, c->vbuf);
  LAUNCHCHK(c);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->ninv_valid = false;
  return KB_OK;
}

int kb_get_sigma(kb_ctx* c, int which, int tp, double* out) {
  if (!c || !out || (which != 0 && which != 1)) return KB_ERR_ARG;
  if (!inwin(c, tp)) return seterr(c, KB_ERR_WINDOW, "sigma: %d outside window [%d,%d]", tp, c->st.tmin, c->st.tmax);
  hipLaunchKernelGGL(gather_sigma, dim3((c->V + 255) / 256), dim3(256), 0, c->stream, c->S, c->N, c->V,
                     slot_of(c, tp), which, c->vbuf);
  LAUNCHCHK(c);
  return lat_download(c, out, c->vbuf);
}

int kb_read_bandwidth(int device, long long bytes, double* gbs) {
  if (!gbs || bytes < (1LL << 24)) return KB_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return KB_ERR_HIP;
  const size_t n2 = (size_t)bytes / 16;
  const int blocks = 8192;
  double *x = nullptr, *out = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  int rc = KB_OK;
  if (hipMalloc((void**)&x, n2 * 16) != hipSuccess || hipMalloc((void**)&out, blocks * sizeof(double)) != hipSuccess ||
      hipMemset(x, 0, n2 * 16) != hipSuccess || hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    rc = KB_ERR_OOM;
  } else {
    for (int w = 0; w < 2; ++w) hipLaunchKernelGGL(read_probe_kernel, dim3(blocks), dim3(256), 0, 0, x, n2, out);
    (void)hipEventRecord(a, 0);
    for (int r = 0; r < 10; ++r) hipLaunchKernelGGL(read_probe_kernel, dim3(blocks), dim3(256), 0, 0, x, n2, out);
    (void)hipEventRecord(b, 0);
    float ms = 0.f;
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
        hipEventElapsedTime(&ms, a, b) != hipSuccess)
      rc = KB_ERR_HIP;
    else
      *gbs = (double)n2 * 16.0 / (ms / 10.0) / 1e6;
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (x) (void)hipFree(x);
  if (out) (void)hipFree(out);
  return rc;
}

int kb_ninv_get(kb_ctx* c, int which, int i, int j, double* out) {
  if (!c || !out || which < 0 || which > 4) return KB_ERR_ARG;
  if (c->p.pexp != KB_PEXP_NINVERSE_NLO) return seterr(c, KB_ERR_STATE, "no 1/N resummation in this context");
  if (which >= 3) {
    if (!inwin(c, j) || c->st.tmax < 3) return seterr(c, KB_ERR_WINDOW, "ring: %d outside window [%d,%d]", j, c->st.tmin, c->st.tmax);
    hipLaunchKernelGGL(gather_sigma, dim3((c->V + 255) / 256), dim3(256), 0, c->stream, c->Iring, c->N, c->V,
                       slot_of(c, j), which == 3 ? 0 : 1, c->vbuf);
    LAUNCHCHK(c);
    return lat_download(c, out, c->vbuf);
  }
  if (!inwin(c, i) || !inwin(c, j) || (which == 2 && i <= j))
    return seterr(c, KB_ERR_WINDOW, "1/N entry (%d,%d) outside window [%d,%d]", i, j, c->st.tmin, c->st.tmax);
  const int si = slot_of(c, i), sj = slot_of(c, j);
  const double* src = which == 0 ? c->PF : which == 1 ? c->PR : c->IR;
  const double sg = which == 1 ? (si >= sj ? 1.0 : -1.0) : 1.0;
  hipLaunchKernelGGL(gather_tri, dim3((c->V + 255) / 256), dim3(256), 0, c->stream, src, c->PNP, c->V,
                     ppair_off(si, sj, c->N), sg, c->vbuf);
  LAUNCHCHK(c);
  return lat_download(c, out, c->vbuf);
}

int kb_check_finite(kb_ctx* c) {
  if (!c) return KB_ERR_ARG;
  HIPCHK(c, hipMemsetAsync(c->flag, 0, sizeof(int), c->stream));
  const size_t count = (size_t)c->NB * (size_t)c->NP * MB * 2;
  hipLaunchKernelGGL(finite_kernel, dim3(4096), dim3(256), 0, c->stream, c->H, count, c->flag);
  LAUNCHCHK(c);
  int f = 0;
  HIPCHK(c, hipMemcpyAsync(&f, c->flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (f) return seterr(c, KB_ERR_NONFINITE, "non-finite value in the two-time window");
  return KB_OK;
}

// ---- checkpoint and restart -------------------------------------------------
namespace {

struct WinHeader {
  char magic[8];  // "QFTKBWIN"
  int version;
  int sym;
  kb_params p;
  int V, N;
  long long NP;
  StepState st;
  double initmass;
  int init_particles;
  int pad;
  unsigned long long hcount, scount, kcount;  // doubles of H, S and kL2 that follow (then 8 of m2)
};
constexpr int WIN_VERSION = 1;
constexpr int WIN_M2 = 8;

void win_counts(const kb_ctx* c, unsigned long long& h, unsigned long long& s, unsigned long long& k) {
  h = (unsigned long long)c->NB * (unsigned long long)c->NP * MB * 2;
  s = (unsigned long long)c->NB * c->N * MB * 2;
  k = (unsigned long long)c->Vpad;
}

unsigned long long win_bytes(const kb_ctx* c) {
  unsigned long long h, s, k;
  win_counts(c, h, s, k);
  return sizeof(WinHeader) + sizeof(double) * (h + s + k + WIN_M2);
}

bool same_params(const kb_params& a, const kb_params& b) {
  return a.Nx == b.Nx && a.sdim == b.sdim && a.NstepsinMemory == b.NstepsinMemory && a.ONgroup == b.ONgroup &&
         a.dt == b.dt && a.Lambda == b.Lambda && a.Mass == b.Mass && a.Mass2 == b.Mass2 && a.pexp == b.pexp;
}

}  // namespace

int kb_window_bytes(const kb_ctx* c, unsigned long long* bytes) {
  if (!c || !bytes) return KB_ERR_ARG;
  *bytes = win_bytes(c);
  return KB_OK;
}

int kb_save_window(kb_ctx* c, void* buf, unsigned long long bytes) {
  if (!c || !buf) return KB_ERR_ARG;
  if (bytes != win_bytes(c)) return seterr(c, KB_ERR_ARG, "save: buffer of %llu bytes, snapshot is %llu", bytes, win_bytes(c));
  if (!c->initialised) return seterr(c, KB_ERR_STATE, "save: context not initialised");
  if (c->adj || c->in_ensemble) return seterr(c, KB_ERR_STATE, "save: not while recording or for ensemble members");
  WinHeader h{};
  std::memcpy(h.magic, "QFTKBWIN", 8);
  h.version = WIN_VERSION;
  h.sym = c->sym ? 1 : 0;
  h.p = c->p;
  h.V = c->V;
  h.N = c->N;
  h.NP = c->NP;
  h.st = c->st;
  h.initmass = c->initmass;
  h.init_particles = c->init_particles ? 1 : 0;
  win_counts(c, h.hcount, h.scount, h.kcount);
  char* out = static_cast<char*>(buf);
  std::memcpy(out, &h, sizeof(h));
  double* d = reinterpret_cast<double*>(out + sizeof(h));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(d, c->H, sizeof(double) * h.hcount, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(d + h.hcount, c->S, sizeof(double) * h.scount, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(d + h.hcount + h.scount, c->kL2, sizeof(double) * h.kcount, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(d + h.hcount + h.scount + h.kcount, c->m2, sizeof(double) * WIN_M2, hipMemcpyDeviceToHost));
  return KB_OK;
}

int kb_load_window(kb_ctx* c, const void* buf, unsigned long long bytes) {
  if (!c || !buf) return KB_ERR_ARG;
  if (c->adj || c->in_ensemble) return seterr(c, KB_ERR_STATE, "load: not while recording or for ensemble members");
  if (bytes != win_bytes(c)) return seterr(c, KB_ERR_ARG, "load: snapshot of %llu bytes, this context takes %llu", bytes, win_bytes(c));
  WinHeader h;
  std::memcpy(&h, buf, sizeof(h));
  unsigned long long hc, sc, kc;
  win_counts(c, hc, sc, kc);
  if (std::memcmp(h.magic, "QFTKBWIN", 8) != 0 || h.version != WIN_VERSION)
    return seterr(c, KB_ERR_ARG, "load: not a window snapshot of this format");
  if (!same_params(h.p, c->p) || h.sym != (c->sym ? 1 : 0) || h.V != c->V || h.N != c->N || h.NP != c->NP ||
      h.hcount != hc || h.scount != sc || h.kcount != kc)
    return seterr(c, KB_ERR_ARG, "load: snapshot taken with other parameters or storage");
  // the window after step tmax is [max(1, tmax - N + 1), tmax]; any other
  // header would make the next step sum slots outside the window
  if (h.st.tmax < 2 || h.st.tmin != (h.st.tmax > c->N ? h.st.tmax - c->N + 1 : 1))
    return seterr(c, KB_ERR_ARG, "load: snapshot window [%d,%d] is not a window of N_mem = %d", h.st.tmin,
                  h.st.tmax, c->N);
  const double* d = reinterpret_cast<const double*>(static_cast<const char*>(buf) + sizeof(h));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(c->H, d, sizeof(double) * hc, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->S, d + hc, sizeof(double) * sc, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->kL2, d + hc + sc, sizeof(double) * kc, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->m2, d + hc + sc + kc, sizeof(double) * WIN_M2, hipMemcpyHostToDevice));
  c->st = state_of_step(h.st.tmax, c->N);
  int rc = upload_state(c);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->initmass = h.initmass;
  c->init_particles = h.init_particles != 0;
  c->initialised = true;
  c->ninv_valid = false;
  drop_graph(c);
  return KB_OK;
}

int kb_fill_free_window(kb_ctx* c, int tmax, const double* n, double mass) {
  if (!c) return KB_ERR_ARG;
  if (tmax < c->N) return seterr(c, KB_ERR_ARG, "fill: tmax=%d < NstepsinMemory=%d", tmax, c->N);
  int rc;
  const double* dn = nullptr;
  if (n) {
    if ((rc = lat_upload(c, c->vbuf, n))) return rc;
    dn = c->vbuf;
  }
  const long total = (long)c->NB * MB * c->N;
  const dim3 grid((unsigned)((total + 255) / 256), (unsigned)c->N);
  hipLaunchKernelGGL(fill_free_kernel, grid, dim3(256), 0, c->stream, c->H, c->NP, c->V, c->N, tmax, c->kL2, dn,
                     mass, c->p.dt, c->NB);
  LAUNCHCHK(c);
  c->st = StepState{};
  c->st.tmin = tmax - c->N + 1;
  c->st.tmax = tmax;
  if ((rc = upload_state(c))) return rc;
  c->initialised = true;
  c->ninv_valid = false;
  if ((rc = launch_hartree(c, tmax, tmax))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return KB_OK;
}

int kb_synchronize(kb_ctx* c) {
  if (!c) return KB_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return KB_OK;
}

int kb_set_stream(kb_ctx* c, void* s) {
  if (!c) return KB_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->stream = s ? (hipStream_t)s : c->own;
  drop_graph(c);
  return KB_OK;
}

int kb_enable_timing(kb_ctx* c, int on) {
  if (!c) return KB_ERR_ARG;
  c->timing = on != 0;
  return KB_OK;
}

int kb_timing(kb_ctx* c, int stage, double* total_ms, long long* count) {
  if (!c || stage < 0 || stage >= NSTAGE) return KB_ERR_ARG;
  int rc = collect_timing(c);
  if (rc) return rc;
  if (total_ms) *total_ms = c->tsum[stage];
  if (count) *count = c->tcount[stage];
  return KB_OK;
}

int kb_tile_shape(int variant, int* rows, int* batch, int* threads) {
  if (variant < 0 || variant >= NMEMINT_VARIANTS) return KB_ERR_ARG;
  if (rows) *rows = kMemintVariants[variant].rows;
  if (batch) *batch = kMemintVariants[variant].batch;
  if (threads) *threads = kMemintVariants[variant].threads;
  return KB_OK;
}

int kb_set_tiles(kb_ctx* c, int variant) {
  if (!c || variant < 0 || variant >= NMEMINT_VARIANTS) return KB_ERR_ARG;
  if (c->in_ensemble) return seterr(c, KB_ERR_STATE, "tiles: ensemble members use the default shape");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->tile = variant;
  const MemintArgs ma = memint_args(c);
  HIPCHK(c, hipMemcpy(c->margs_dev, &ma, sizeof(MemintArgs), hipMemcpyHostToDevice));
  drop_graph(c);
  return KB_OK;
}

int kb_get_tiles(const kb_ctx* c, int* variant) {
  if (!c || !variant) return KB_ERR_ARG;
  *variant = c->tile;
  return KB_OK;
}

int kb_autotune(kb_ctx* c, int reps, int* best, double* ms) {
  if (!c || reps < 1) return KB_ERR_ARG;
  if (!c->initialised || c->st.tmax < 3)
    return seterr(c, KB_ERR_STATE, "autotune: take a step first (the last step's integral is re-timed)");
  if (c->in_ensemble || c->adj) return seterr(c, KB_ERR_STATE, "autotune: not for ensemble members or while recording");
  std::vector<double> t(NMEMINT_VARIANTS, 0.0);
  hipEvent_t a = take_event(c), b = take_event(c);
  if (!a || !b) return seterr(c, KB_ERR_HIP, "autotune: no events");
  int rc = KB_OK;
  for (int v = 0; v < NMEMINT_VARIANTS && rc == KB_OK; ++v) {
    if ((rc = kb_set_tiles(c, v))) break;
    // re-evaluates row t of the last step from the same inputs, in the form
    // the steps use (split in orbit storage)
    for (int k = 0; k <= reps && rc == KB_OK; ++k) {
      if (k == 1 && hipEventRecord(a, c->stream) != hipSuccess) rc = seterr(c, KB_ERR_HIP, "autotune: event");
      if (c->msplit > 1)
        rc = launch_memint_split(c);
      else if (launch_memint(v, c->NB, memint_lds(c->N), c->stream, c->margs_dev, c->wg0_dev, 1) != hipSuccess)
        rc = seterr(c, KB_ERR_HIP, "autotune: launch failed");
    }
    if (rc) break;
    float e_ms = 0.f;
    if (hipEventRecord(b, c->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
        hipEventElapsedTime(&e_ms, a, b) != hipSuccess)
      rc = seterr(c, KB_ERR_HIP, "autotune: timing failed");
    t[v] = e_ms / reps;
  }
  c->freeEv.push_back(a);
  c->freeEv.push_back(b);
  if (rc) return rc;
  int bv = 0;
  for (int v = 1; v < NMEMINT_VARIANTS; ++v)
    if (t[v] < t[bv]) bv = v;
  if ((rc = kb_set_tiles(c, bv))) return rc;
  if (best) *best = bv;
  if (ms)
    for (int v = 0; v < NMEMINT_VARIANTS; ++v) ms[v] = t[v];
  return KB_OK;
}

int kb_fft_shape(int variant, int* threads) {
  if (variant < 0 || variant >= NFFT_VARIANTS) return KB_ERR_ARG;
  if (threads) *threads = kFFTThreads[variant];
  return KB_OK;
}

int kb_set_fft_tiles(kb_ctx* c, int variant) {
  if (!c || (variant != KB_FFT_ORBIT_PRODUCT && (variant < 0 || variant >= NFFT_VARIANTS)) ||
      (variant == KB_FFT_ORBIT_PRODUCT && !c->Mo))
    return KB_ERR_ARG;
  // orbit storage past the plane kernels' sizes has only the orbit product
  if (variant != KB_FFT_ORBIT_PRODUCT && c->sym && !sigma_planes_supported(c->p.Nx, c->p.sdim)) return KB_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (variant != KB_FFT_ORBIT_PRODUCT) {
    int rc = ensure_plane_workspace(c);
    if (rc) return rc;
  }
  c->fft_tile = variant;
  drop_graph(c);
  return KB_OK;
}

int kb_get_fft_tiles(const kb_ctx* c, int* variant) {
  if (!c || !variant) return KB_ERR_ARG;
  *variant = c->fft_tile;
  return KB_OK;
}

int kb_autotune_fft(kb_ctx* c, int reps, int* best, double* ms) {
  if (!c || reps < 1) return KB_ERR_ARG;
  if (!c->initialised || c->st.tmax < 3)
    return seterr(c, KB_ERR_STATE, "autotune: take a step first (the last step's self-energy is re-timed)");
  if (c->in_ensemble || c->adj) return seterr(c, KB_ERR_STATE, "autotune: not for ensemble members or while recording");
  if (!sunset_expansion(c->p.pexp) || !sigma_planes_supported(c->p.Nx, c->p.sdim))
    return seterr(c, KB_ERR_ARG, "autotune: no plane self-energy stage for this context");
  // the plane kernels' block sizes, then (orbit storage) the orbit product
  const int nv = NFFT_VARIANTS + (c->Mo ? 1 : 0);
  if (int rw = ensure_plane_workspace(c)) return rw;
  std::vector<double> t(nv, 0.0);
  hipEvent_t a = take_event(c), b = take_event(c);
  if (!a || !b) return seterr(c, KB_ERR_HIP, "autotune: no events");
  int rc = KB_OK;
  for (int v = 0; v < nv && rc == KB_OK; ++v) {
    c->fft_tile = v < NFFT_VARIANTS ? v : KB_FFT_ORBIT_PRODUCT;
    // recomputes Sigma(t-1, .) of the last step: the same values
    for (int k = 0; k <= reps && rc == KB_OK; ++k) {
      if (k == 1 && hipEventRecord(a, c->stream) != hipSuccess) rc = seterr(c, KB_ERR_HIP, "autotune: event");
      if (rc == KB_OK) rc = launch_sigma(c);
    }
    if (rc) break;
    float e_ms = 0.f;
    if (hipEventRecord(b, c->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
        hipEventElapsedTime(&e_ms, a, b) != hipSuccess)
      rc = seterr(c, KB_ERR_HIP, "autotune: timing failed");
    t[v] = e_ms / reps;
  }
  c->freeEv.push_back(a);
  c->freeEv.push_back(b);
  if (rc) return rc;
  int bv = 0;
  for (int v = 1; v < nv; ++v)
    if (t[v] < t[bv]) bv = v;
  if (bv == NFFT_VARIANTS) bv = KB_FFT_ORBIT_PRODUCT;
  if ((rc = kb_set_fft_tiles(c, bv))) return rc;
  if (best) *best = bv;
  if (ms)
    for (int v = 0; v < nv; ++v) ms[v] = t[v];
  return KB_OK;
}

int kb_reset_timing(kb_ctx* c) {
  if (!c) return KB_ERR_ARG;
  int rc = collect_timing(c);
  if (rc) return rc;
  for (int s = 0; s < NSTAGE; ++s) {
    c->tsum[s] = 0;
    c->tcount[s] = 0;
  }
  return KB_OK;
}

}  // extern "C"

// ===========================================================================
// ensembles
// ===========================================================================
struct kb_ens {
  std::vector<kb_ctx*> m;
  int device = 0;
  hipStream_t stream = nullptr;
  // members' self-energy stages run concurrently on up to 4 lanes: lane 0
  // is stream itself, lanes 1.. are side streams forked from and joined back
  // into it (4 lanes on 4 hardware queues; member i on lane[i])
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> join;
  std::vector<int> lane;
  hipEvent_t fork = nullptr;
  StepState** stptr_dev = nullptr;
  int* N_dev = nullptr;
  // memory-integral argument tables, runs ordered by lane: lane j owns runs
  // [lrun[j], lrun[j+1]) and lwg[j] workgroups; wg_dev holds each run's
  // first workgroup over all lanes (one pooled launch), lwg_dev within its
  // lane (one launch per lane)
  MemintArgs* margs_dev = nullptr;
  int* wg_dev = nullptr;
  int* lwg_dev = nullptr;
  std::vector<int> lrun, lwg, lmaxN;
  int total_wg = 0;
  int maxN = 0;
  // a memory integral per lane after the lane's stages (every member in
  // orbit storage: its stage is matrix products that the other lanes'
  // integrals overlap), else one pooled launch after all stages
  bool lane_integrals = false;
  int tile = 0;     // shape of the pooled memory-integral launch (kMemintVariants)
  int last_t = 0;   // t of the last ensemble step (0: none yet)
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool timing = false;
  std::vector<EventPair> pending;
  double tsum = 0.0;
  long long tcount = 0;
  std::string err;
};

namespace {

int ens_err(kb_ens* e, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (e) e->err = buf;
  return code;
}

void ens_drop_graph(kb_ens* e) {
  if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->gexec = nullptr;
  e->graph = nullptr;
}

void ens_free(kb_ens* e) {
  ens_drop_graph(e);
  for (auto* c : e->m) {
    if (c) {
      c->stream = c->own;
      kb_destroy(c);
    }
  }
  for (void* p : {(void*)e->stptr_dev, (void*)e->N_dev, (void*)e->margs_dev, (void*)e->wg_dev, (void*)e->lwg_dev})
    if (p) (void)hipFree(p);
  for (auto& ev : e->pending) {
    (void)hipEventDestroy(ev.a);
    (void)hipEventDestroy(ev.b);
  }
  for (auto ev : e->join)
    if (ev) (void)hipEventDestroy(ev);
  if (e->fork) (void)hipEventDestroy(e->fork);
  for (size_t j = 1; j < e->side.size(); ++j)
    if (e->side[j]) (void)hipStreamDestroy(e->side[j]);
  if (e->stream) (void)hipStreamDestroy(e->stream);
}

// the members' stages on the side streams: each member to the least-loaded
// stream, largest first; the load of a member is the flops of its transforms
int ens_make_lanes(kb_ens* e) {
  const int nr = (int)e->m.size();
  const int ns = std::min(nr, 4);
  e->lane.assign(nr, 0);
  if (ns < 2) return KB_OK;
  e->side.assign(ns, nullptr);
  e->join.assign(ns, nullptr);
  e->side[0] = e->stream;
  for (int j = 1; j < ns; ++j)
    if (hipStreamCreateWithFlags(&e->side[j], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->join[j], hipEventDisableTiming) != hipSuccess)
      return KB_ERR_HIP;
  if (hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess) return KB_ERR_HIP;
  std::vector<double> work(nr);
  std::vector<int> order(nr);
  for (int i = 0; i < nr; ++i) {
    const kb_ctx* c = e->m[i];
    const double sites = c->Mo ? (double)c->Mp * c->Mp : 10.0 * c->p.Nx * std::pow((double)c->p.Nx, c->p.sdim);
    work[i] = sites * c->N;
    order[i] = i;
  }
  std::sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
  std::vector<double> load(ns, 0.0);
  for (int i : order) {
    const int j = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    e->lane[i] = j;
    load[j] += work[i];
  }
  return KB_OK;
}

// one ensemble step: advance all windows; each member's self-energy and
// Hartree kernels on its lane; the memory integrals of all members in one
// launch after the lanes have joined, or (lane_integrals) one per lane
// after that lane's stages
int ens_enqueue(kb_ens* e) {
  const int nr = (int)e->m.size();
  hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(64 * ((nr + 63) / 64)), 0, e->stream, e->stptr_dev, e->N_dev, nr);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return ens_err(e, KB_ERR_HIP, "advance launch failed: %s", hipGetErrorString(err));
  const bool forked = !e->side.empty();
  if (forked) {
    err = hipEventRecord(e->fork, e->stream);
    for (size_t j = 1; j < e->side.size() && err == hipSuccess; ++j) err = hipStreamWaitEvent(e->side[j], e->fork, 0);
    if (err != hipSuccess) return ens_err(e, KB_ERR_HIP, "fork: %s", hipGetErrorString(err));
  }
  // per lane: its members' self-energy and Hartree kernels, then (with
  // lane_integrals) the memory integral over its members
  int rc = KB_OK;
  const int nl = (int)e->lwg.size();
  for (int j = 0; j < nl && rc == KB_OK; ++j) {
    hipStream_t sj = forked ? e->side[j] : e->stream;
    for (size_t i = 0; i < e->m.size() && rc == KB_OK; ++i) {
      if (e->lane[i] != j) continue;
      kb_ctx* c = e->m[i];
      c->stream = sj;
      if (sunset_expansion(c->p.pexp) && (rc = launch_sigma(c)))
        rc = ens_err(e, rc, "member self-energy: %s", c->err.c_str());
      else if (c->p.pexp == KB_PEXP_NINVERSE_NLO && (rc = launch_ninv_stage(c)))
        rc = ens_err(e, rc, "member 1/N self-energy: %s", c->err.c_str());
      else if ((rc = launch_hartree_step(c)))
        rc = ens_err(e, rc, "member Hartree: %s", c->err.c_str());
    }
    if (rc == KB_OK && e->lane_integrals && e->lwg[j] > 0) {
      // one shape for all members (their item tables follow its row height)
      err = launch_memint(e->tile, e->lwg[j], memint_lds(e->lmaxN[j]), sj, e->margs_dev + e->lrun[j],
                          e->lwg_dev + e->lrun[j], e->lrun[j + 1] - e->lrun[j]);
      if (err != hipSuccess) rc = ens_err(e, KB_ERR_HIP, "memint launch failed: %s", hipGetErrorString(err));
    }
  }
  for (auto* c : e->m) c->stream = e->stream;
  if (forked) {
    // joined even after a failed launch, so that a capture ends cleanly
    for (size_t j = 1; j < e->side.size(); ++j) {
      hipError_t je = hipEventRecord(e->join[j], e->side[j]);
      if (je == hipSuccess) je = hipStreamWaitEvent(e->stream, e->join[j], 0);
      if (je != hipSuccess && rc == KB_OK) rc = ens_err(e, KB_ERR_HIP, "join: %s", hipGetErrorString(je));
    }
  }
  if (rc) return rc;
  if (!e->lane_integrals) {
    err = launch_memint(e->tile, e->total_wg, memint_lds(e->maxN), e->stream, e->margs_dev, e->wg_dev, nr);
    if (err != hipSuccess) return ens_err(e, KB_ERR_HIP, "memint launch failed: %s", hipGetErrorString(err));
  }
  return KB_OK;
}

// the memory-integral argument tables in lane order; a member's tile
// selects its item table
int ens_upload_tables(kb_ens* e) {
  const int nl = std::max(1, (int)e->side.size());
  std::vector<MemintArgs> ma;
  std::vector<int> wg, lwg;
  e->lrun.assign(nl + 1, 0);
  e->lwg.assign(nl, 0);
  e->lmaxN.assign(nl, 0);
  e->maxN = 0;
  int total = 0;
  for (int j = 0; j < nl; ++j) {
    e->lrun[j] = (int)ma.size();
    for (size_t i = 0; i < e->m.size(); ++i) {
      const kb_ctx* c = e->m[i];
      if (e->lane[i] != j) continue;
      ma.push_back(memint_args(c));
      wg.push_back(total);
      lwg.push_back(e->lwg[j]);
      total += c->NB;
      e->lwg[j] += c->NB;
      e->lmaxN[j] = std::max(e->lmaxN[j], c->N);
      e->maxN = std::max(e->maxN, c->N);
    }
  }
  e->lrun[nl] = (int)ma.size();
  // the pooled launch has one workgroup per site block of every member: keep
  // its work-items (up to 1024 per workgroup) below 2^32
  if ((long)total > (long)(0xffffffffL / 1024))
    return ens_err(e, KB_ERR_ARG, "ensemble: %d site blocks in all, at most %ld", total, 0xffffffffL / 1024);
  e->total_wg = total;
  const size_t n = ma.size();
  if (!e->margs_dev && (hipMalloc((void**)&e->margs_dev, sizeof(MemintArgs) * n) != hipSuccess ||
                        hipMalloc((void**)&e->wg_dev, sizeof(int) * n) != hipSuccess ||
                        hipMalloc((void**)&e->lwg_dev, sizeof(int) * n) != hipSuccess))
    return ens_err(e, KB_ERR_OOM, "argument tables: out of device memory");
  if (hipMemcpy(e->margs_dev, ma.data(), sizeof(MemintArgs) * n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->wg_dev, wg.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->lwg_dev, lwg.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess)
    return ens_err(e, KB_ERR_HIP, "argument tables: copy failed");
  return KB_OK;
}

// the memory-integral shape, one for all members: every member's work items
// are cut to its row height
int ens_set_tile(kb_ens* e, int v) {
  hipError_t err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) return ens_err(e, KB_ERR_HIP, "tiles: %s", hipGetErrorString(err));
  for (auto* c : e->m) c->tile = v;
  int rc = ens_upload_tables(e);
  if (rc) return rc;
  e->tile = v;
  ens_drop_graph(e);
  return KB_OK;
}

}  // namespace

extern "C" {

const char* kb_ens_last_error(const kb_ens* e) { return e ? e->err.c_str() : "null ensemble"; }

int kb_ens_set_tiles(kb_ens* e, int variant) {
  if (!e || variant < 0 || variant >= NMEMINT_VARIANTS) return KB_ERR_ARG;
  (void)hipSetDevice(e->device);
  return ens_set_tile(e, variant);
}

int kb_ens_get_tiles(const kb_ens* e, int* variant) {
  if (!e || !variant) return KB_ERR_ARG;
  *variant = e->tile;
  return KB_OK;
}

int kb_ens_autotune(kb_ens* e, int reps, int* best, double* ms) {
  if (!e || reps < 1) return KB_ERR_ARG;
  if (!e->last_t) return ens_err(e, KB_ERR_STATE, "autotune: take an ensemble step first (its integral is re-timed)");
  for (size_t i = 0; i < e->m.size(); ++i) {
    const kb_ctx* c = e->m[i];
    if (!c->initialised || c->st.tmax != e->last_t)
      return ens_err(e, KB_ERR_STATE, "autotune: member %zu has changed since the last ensemble step", i);
  }
  (void)hipSetDevice(e->device);
  const int nr = (int)e->m.size();
  std::vector<double> t(NMEMINT_VARIANTS, 0.0);
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    if (a) (void)hipEventDestroy(a);
    return ens_err(e, KB_ERR_HIP, "autotune: no events");
  }
  int rc = KB_OK;
  for (int v = 0; v < NMEMINT_VARIANTS && rc == KB_OK; ++v) {
    if ((rc = ens_set_tile(e, v))) break;
    // re-evaluates row t of the last step of every member from the same inputs
    for (int k = 0; k <= reps && rc == KB_OK; ++k) {
      if (k == 1 && hipEventRecord(a, e->stream) != hipSuccess) rc = ens_err(e, KB_ERR_HIP, "autotune: event");
      if (rc == KB_OK &&
          launch_memint(v, e->total_wg, memint_lds(e->maxN), e->stream, e->margs_dev, e->wg_dev, nr) != hipSuccess)
        rc = ens_err(e, KB_ERR_HIP, "autotune: launch failed");
    }
    if (rc) break;
    float e_ms = 0.f;
    if (hipEventRecord(b, e->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
        hipEventElapsedTime(&e_ms, a, b) != hipSuccess)
      rc = ens_err(e, KB_ERR_HIP, "autotune: timing failed");
    t[v] = e_ms / reps;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) return rc;
  int bv = 0;
  for (int v = 1; v < NMEMINT_VARIANTS; ++v)
    if (t[v] < t[bv]) bv = v;
  if ((rc = ens_set_tile(e, bv))) return rc;
  if (best) *best = bv;
  if (ms)
    for (int v = 0; v < NMEMINT_VARIANTS; ++v) ms[v] = t[v];
  return KB_OK;
}

int kb_ens_create(const kb_params* params, int nruns, kb_ens** out) {
  std::vector<int> full(nruns > 0 ? nruns : 0, 0);
  return kb_ens_create_sym(params, full.data(), nruns, out);
}

int kb_ens_create_sym(const kb_params* params, const int* symmetric, int nruns, kb_ens** out) {
  if (!params || nruns < 1 || nruns > 64 || !out) return KB_ERR_ARG;
  *out = nullptr;
  kb_ens* e = new (std::nothrow) kb_ens();
  if (!e) return KB_ERR_OOM;
  e->device = params[0].device;
  auto fail = [&](int code) {
    ens_free(e);
    delete e;
    return code;
  };
  for (int i = 0; i < nruns; ++i) {
    if (params[i].device != e->device) return fail(KB_ERR_ARG);
    kb_ctx* c = nullptr;
    // an orbit-storage member keeps its own storage: the memory-integral
    // arguments carry its stored sites, its self-energy stage is its own
    int rc = create_ctx(&params[i], !symmetric || symmetric[i] != 0, &c);
    if (rc) return fail(rc);
    e->m.push_back(c);
  }
  if (hipSetDevice(e->device) != hipSuccess) return fail(KB_ERR_HIP);
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) return fail(KB_ERR_HIP);
  if (ens_make_lanes(e) != KB_OK) return fail(KB_ERR_HIP);
  std::vector<StepState*> sp;
  std::vector<int> Ns;
  for (auto* c : e->m) {
    (void)hipStreamSynchronize(c->stream);
    c->stream = e->stream;
    c->in_ensemble = true;
    sp.push_back(c->st_dev);
    Ns.push_back(c->N);
  }
  if (hipMalloc((void**)&e->stptr_dev, sizeof(StepState*) * nruns) != hipSuccess ||
      hipMalloc((void**)&e->N_dev, sizeof(int) * nruns) != hipSuccess)
    return fail(KB_ERR_OOM);
  if (hipMemcpy(e->stptr_dev, sp.data(), sizeof(StepState*) * nruns, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->N_dev, Ns.data(), sizeof(int) * nruns, hipMemcpyHostToDevice) != hipSuccess)
    return fail(KB_ERR_HIP);
  if (int rc = ens_upload_tables(e)) return fail(rc);
  e->lane_integrals = e->side.size() > 1;
  for (auto* c : e->m) e->lane_integrals = e->lane_integrals && c->sym;
  *out = e;
  return KB_OK;
}

int kb_ens_destroy(kb_ens* e) {
  if (!e) return KB_OK;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  ens_free(e);
  delete e;
  return KB_OK;
}

int kb_ens_member(kb_ens* e, int i, kb_ctx** member) {
  if (!e || !member || i < 0 || i >= (int)e->m.size()) return KB_ERR_ARG;
  *member = e->m[i];
  return KB_OK;
}

int kb_ens_step(kb_ens* e, int t) {
  if (!e) return KB_ERR_ARG;
  for (size_t i = 0; i < e->m.size(); ++i) {
    const kb_ctx* c = e->m[i];
    if (!c->initialised) return ens_err(e, KB_ERR_STATE, "member %zu not initialised", i);
    if (t != c->st.tmax + 1 || t < 3)
      return ens_err(e, KB_ERR_WINDOW, "member %zu: t=%d but window is [%d,%d]", i, t, c->st.tmin, c->st.tmax);
  }
  for (auto* c : e->m) expand(c);
  for (auto* c : e->m)
    if (c->p.pexp == KB_PEXP_NINVERSE_NLO && !c->ninv_valid) {
      const int rc = ninv_rebuild(c);  // eager, before the step's launches or graph
      if (rc) return ens_err(e, rc, "member 1/N rebuild: %s", c->err.c_str());
    }
  EventPair ev{nullptr, nullptr};
  if (e->timing) {
    if (hipEventCreate(&ev.a) != hipSuccess || hipEventCreate(&ev.b) != hipSuccess)
      return ens_err(e, KB_ERR_HIP, "event creation failed");
    (void)hipEventRecord(ev.a, e->stream);
  }
  int rc = KB_OK;
  if (e->use_graph) {
    if (!e->gexec) {
      hipError_t err = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
      if (err != hipSuccess) return ens_err(e, KB_ERR_HIP, "capture: %s", hipGetErrorString(err));
      rc = ens_enqueue(e);
      hipGraph_t g = nullptr;
      err = hipStreamEndCapture(e->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (err != hipSuccess) return ens_err(e, KB_ERR_HIP, "capture: %s", hipGetErrorString(err));
      e->graph = g;
      err = hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0);
      if (err != hipSuccess) return ens_err(e, KB_ERR_HIP, "instantiate: %s", hipGetErrorString(err));
    }
    hipError_t err = hipGraphLaunch(e->gexec, e->stream);
    if (err != hipSuccess) return ens_err(e, KB_ERR_HIP, "graph launch: %s", hipGetErrorString(err));
  } else {
    rc = ens_enqueue(e);
    if (rc) return rc;
  }
  if (e->timing) {
    (void)hipEventRecord(ev.b, e->stream);
    e->pending.push_back(ev);
  }
  e->last_t = t;
  return KB_OK;
}

int kb_ens_use_graph(kb_ens* e, int on) {
  if (!e) return KB_ERR_ARG;
  if (!on) ens_drop_graph(e);
  e->use_graph = on != 0;
  return KB_OK;
}

int kb_ens_synchronize(kb_ens* e) {
  if (!e) return KB_ERR_ARG;
  hipError_t err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) return ens_err(e, KB_ERR_HIP, "synchronize: %s", hipGetErrorString(err));
  return KB_OK;
}

int kb_ens_enable_timing(kb_ens* e, int on) {
  if (!e) return KB_ERR_ARG;
  e->timing = on != 0;
  return KB_OK;
}

int kb_ens_timing(kb_ens* e, double* total_ms, long long* count) {
  if (!e) return KB_ERR_ARG;
  hipError_t err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) return ens_err(e, KB_ERR_HIP, "synchronize: %s", hipGetErrorString(err));
  for (auto& ev : e->pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev.a, ev.b);
    e->tsum += ms;
    e->tcount += 1;
    (void)hipEventDestroy(ev.a);
    (void)hipEventDestroy(ev.b);
  }
  e->pending.clear();
  if (total_ms) *total_ms = e->tsum;
  if (count) *count = e->tcount;
  return KB_OK;
}

}  // extern "C"
