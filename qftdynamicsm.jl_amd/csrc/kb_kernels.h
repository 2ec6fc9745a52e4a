// kb_kernels.h -- kernel argument blocks and declarations (libqftkb).
#pragma once
#include <hip/hip_runtime.h>

#include "kb_device.h"

// kernels and launch helpers are internal to the library
#pragma GCC visibility push(hidden)

namespace qftkb {

constexpr int FFT_MAXF = 24;

struct FFTPlan {
  int n;
  int nf;
  int fac[FFT_MAXF];
};

enum SigmaFlags {
  SIGMA_GATHER = 1,   // read F(t-1,t'), rho(t-1,t') from the history (axis 0, forward)
  SIGMA_SUNSET = 2,   // after the forward transform: pointwise sunset, inverse transform
  SIGMA_SCATTER = 4,  // write SigF/Sigr to the self-energy ring (axis 0, inverse)
  SIGMA_INVERSE = 8,  // transform direction of a plain pass
};

// Window of the step being taken, after expandSimData! (SimDataStruct.jl:
// 262-278).  Kept in device memory and advanced by advance_kernel, so a
// step's kernels have step-invariant arguments and can be replayed from a
// captured hipGraph.
struct StepState {
  int t, tmin, tmax;
  int L;     // live times t' in [tmin, t-1]: t - tmin
  int Rmax;  // slots [0, Rmax) are in use
  int dead;  // slot overwritten by row t when the window is full, else -1
  int s_t, s_tm1, s_tm2, s_min;
};

// expand the window by one step and derive the slots of step t = tmax + 1
__host__ __device__ inline void advance_state(StepState& s, int N) {
  if (s.tmax + 1 > N) {
    s.tmin += 1;
    s.tmax += 1;
  } else {
    s.tmin = 1;
    s.tmax += 1;
  }
  const int t = s.tmax;
  const bool full = t > N;
  s.t = t;
  s.L = t - s.tmin;
  s.Rmax = full ? N : t - 1;
  s.s_t = (t - 1) % N;
  s.dead = full ? s.s_t : -1;
  s.s_tm1 = (t - 2) % N;
  s.s_tm2 = (t - 3) % N;
  s.s_min = (s.tmin - 1) % N;
}

// the state advance_state reaches for step t, from t alone (used to walk
// the window backwards in the adjoint sweep)
__host__ __device__ inline StepState state_of_step(int t, int N) {
  StepState s{};
  s.tmax = t;
  s.tmin = t > N ? t - N + 1 : 1;
  const bool full = t > N;
  s.t = t;
  s.L = t - s.tmin;
  s.Rmax = full ? N : t - 1;
  s.s_t = (t - 1) % N;
  s.dead = full ? s.s_t : -1;
  s.s_tm1 = (t - 2) % N;
  s.s_tm2 = (t - 3) % N;
  s.s_min = (s.tmin - 1) % N;
  return s;
}

// modes of the plane kernels (kb_sigma.hip)
// SG_PLAIN is the sunset of caller-supplied lattices: rows q < rows of the
// physical F and rho in Fin/Rin (q * V + site), Sigma into SigFout/Sigrout
enum SigmaMode { SG_SUNSET = 0, SG_SUNSET_ADJ = 1, SG_BUBBLE = 2, SG_CHAIN = 3, SG_PLAIN = 4 };

struct SigmaArgs {
  const double* H;
  double* S;
  double2* W;  // workspace rows [N-1][V] complex
  const double2* tw;
  FFTPlan plan;
  long NP;
  int N, V, n, nx, ny, nz;
  int axis, flags, P, npencils;
  const StepState* st;  // rows q < st->L (t' = tmin + q) are live
  double preF, preR;
  double2* X;           // forward: if set, receives (Fx, rx) per row and site; adjoint: read
  const double* Sbar;   // adjoint input (SigF_bar, Sigr_bar), self-energy ring layout
  double* Hbar;         // adjoint output: history-layout adjoint, row t-1 is added to
  // orbit storage (NULL: every lattice site is stored): lattice site ->
  // stored site, and lattice site -> stored site where it is the
  // orbit's representative, else -1
  const int* smap;
  const int* rmap;
  // 1/N expansion: bubble triangles written by SG_BUBBLE, (IF, Ir) ring read
  // by SG_CHAIN, 1/V
  double* PF;
  double* PR;
  const double* Iring;
  long PNP;  // slot pairs per site block of the bubble triangles
  double ivol;
  // SG_PLAIN: input and output lattices, and their number
  const double* Fin;
  const double* Rin;
  double* SigFout;
  double* Sigrout;
  int rows;
  int classical;  // the sunset without its rho^2 terms (LambdaNLOclassical)
};

// The resummation sweeps of the 1/N expansion (kb_ninverse.hip): one
// workgroup per site block; st->t - 1 is the time whose row of Ir and IF is
// computed.
struct NinvArgs {
  const double* PF;  // bubble triangles
  const double* PR;
  double* IR;        // resummed-kernel triangle Ir(later, earlier)
  double* Iring;     // (IF, Ir)(t-1, t') per slot, self-energy layout
  const int2* items;  // (column block of PCB slots, row chunk) work items
  int nitems;
  long PNP;
  int N, V;
  double dt, lam;
  int passes;        // NINV_PASS_* bits
  const StepState* st;
};
enum NinvPasses { NINV_PASS_IR = 1, NINV_PASS_ALL = 7 };
// tile of the resummation sweeps: threads per workgroup, slot rows per work
// item (build-time; the defaults were picked by timing at 32^3, N_mem = 400)
#ifndef KB_NINV_THREADS
#define KB_NINV_THREADS 512
#endif
#ifndef KB_NINV_ROWS
#define KB_NINV_ROWS 128
#endif
constexpr int NINV_THREADS = KB_NINV_THREADS;
constexpr int NINV_ROWS = KB_NINV_ROWS;
__global__ void ninv_resum_kernel(NinvArgs a);
// split form for few site blocks (cubic-orbit storage): per sweep, nsplit
// workgroups per site block write partial accumulators (N MB doubles each)
// to part; the sweep's finish kernel sums them, runs the phase that follows
// and passes the next coefficient vector through cvg ([site block][slot][lane])
template <int SWEEP>
__global__ void ninv_split_sweep(NinvArgs a, int nsplit, double* part, const double* cvg);
template <int SWEEP>
__global__ void ninv_split_finish(NinvArgs a, int nsplit, const double* part, double* cvg);
constexpr int NINV_FINISH_THREADS = 1024;
inline size_t ninv_lds(int N) { return sizeof(double) * 2 * (size_t)N * MB; }

// Per-run arguments of memint_step; step-dependent values come from *st.
struct MemintArgs {
  double* H;
  const double* S;
  const double* kL2;
  const double* m2;
  long NP;
  int N, V;
  double dt;
  int integral;       // 0 for LambdaLO
  const int2* items;  // (column block J, row chunk k) in storage order
  int nitems;
  const StepState* st;
};

struct MeasArgs {
  const double* wgt;  // sites each stored site stands for (orbit storage), NULL = 1
  const double* H;
  const double* S;
  const double* kL2;
  long NP;
  int N, V;
  int s_t, s_tm1;
  int tmin, t;
  double dt;
  int integral;
  double *outF, *outddF, *outr, *outn, *outom, *outc;
  double* partial;
};

__global__ void sigma_pass(SigmaArgs a);
// wgt: sites each stored site stands for (orbit storage), NULL = 1
__global__ void hartree_kernel(const double* H, long NP, int V, long pair, double hfac, double Mass2,
                               const double* wgt, double* out);
// Hartree mass of the step being taken: F(t-1,t-1) from *st; one workgroup
// of HARTREE_THREADS per HARTREE_SITES sites, part: hartree_groups(V)
// partial sums followed by a zeroed counter (hartree_part_doubles(V) doubles);
// sum_out (may be null) receives the weighted sum of F(t-1,t-1) itself
constexpr int HARTREE_THREADS = 256;
constexpr int HARTREE_LOADS = 8;
constexpr int HARTREE_SITES = HARTREE_THREADS * HARTREE_LOADS;
inline int hartree_groups(int V) { return (V + HARTREE_SITES - 1) / HARTREE_SITES; }
inline size_t hartree_part_doubles(int V) { return (size_t)hartree_groups(V) + 2; }
__global__ void hartree_step_kernel(const double* H, long NP, int V, int N, const StepState* st, double hfac,
                                    double Mass2, const double* wgt, double* out, double* part, double* sum_out);
// one launch over all runs: workgroups [wg_start[r], wg_start[r+1]) belong to run r
template <int ROWS, int B, int TPB>
__global__ void memint_step(const MemintArgs* runs, const int* wg_start, int nruns);
// split form for few site blocks: nsplit workgroups per site block write
// partial accumulators (2 N MB doubles each) to part; memint_finish sums
// them and runs the leapfrog (one run, grid = site blocks)
template <int ROWS, int B, int TPB>
__global__ void memint_partial(const MemintArgs* runs, int nsplit, double* part);
__global__ void memint_finish(const MemintArgs* runs, int nsplit, const double* part);

// memory-integral tile shapes: ROWS slot rows per work item, B rows per
// batch (two batches in flight), threads per workgroup; fn is the kernel
// with one workgroup per site block, partial the split form's partial sweep
struct MemintVariant {
  int rows, batch, threads;
  const void* fn;
  const void* partial;
};
constexpr int NMEMINT_VARIANTS = 13;
extern const MemintVariant kMemintVariants[NMEMINT_VARIANTS];
constexpr int MEMINT_ROW_CHOICES[3] = {32, 64, 128};  // item tables kept per context
__global__ void advance_kernel(StepState* const* states, const int* Ns, int nruns);
__global__ void measure_kernel(MeasArgs a);
__global__ void history_kernel(const double* H, const double* S, long NP, int N, int tmin, int t, double* out);
__global__ void gather_pair(const double* H, long NP, int V, long pair, int comp, double sign, double* out);
__global__ void scatter_pair(double* H, long NP, int V, long pair, int comp, double sign, const double* in);
__global__ void gather_sigma(const double* S, int N, int V, int slot, int comp, double* out);
__global__ void gather_tri(const double* P, long PNP, int V, long pp, double sign, double* out);
__global__ void read_probe_kernel(const double* x, size_t n2, double* out);
__global__ void init_rows_kernel(double* H, long NP, int V, int N, const double* F11, const double* F21,
                                 const double* F22, double dt);
__global__ void init_particles_kernel(const double* kL2, const double* n, double m, double dt, int V, double* F11,
                                      double* F21, double* F22);
__global__ void fill_free_kernel(double* H, long NP, int V, int N, int tmax, const double* kL2, const double* n,
                                 double mass, double dt, int nb);
__global__ void finite_kernel(const double* H, size_t count, int* flag);

// self-energy stage on whole planes (kb_sigma.hip), Nx <= 64
// the plane kernels take Nx <= 64, and 1D rows up to 2048 points (one line in LDS)
bool sigma_planes_supported(int Nx, int sdim);
hipError_t sigma_planes_prepare(int n);
// variant: index into kFFTThreads (threads per workgroup of the plane kernels)
constexpr int NFFT_VARIANTS = 4;
extern const int kFFTThreads[NFFT_VARIANTS];
hipError_t launch_sigma_planes(const SigmaArgs& a, int rows, hipStream_t st, int variant = 0);
hipError_t launch_sigma_planes_adjoint(const SigmaArgs& a, int rows, hipStream_t st);
// the sunset of a.rows caller lattices (SG_PLAIN), default block size
hipError_t launch_sigma_planes_plain(const SigmaArgs& a, hipStream_t st);
hipError_t launch_bubble_planes(const SigmaArgs& a, int rows, hipStream_t st);
hipError_t launch_chain_planes(const SigmaArgs& a, int rows, hipStream_t st);
// the same passes on 2D and 3D lattices past the plane kernels (pencils of
// whole rows in W)
hipError_t launch_bubble_pencils(const SigmaArgs& a, int rows, hipStream_t st);
hipError_t launch_chain_pencils(const SigmaArgs& a, int rows, hipStream_t st);
// the sunset keeping its transforms in a.X, and its adjoint, as pencil passes
hipError_t launch_sunset_pencils(const SigmaArgs& a, int rows, hipStream_t st);
hipError_t launch_sunset_adjoint_pencils(const SigmaArgs& a, int rows, hipStream_t st);

// self-energy stage of a cubic-orbit-stored window as two FP64 matrix
// products with the orbit cosine matrix Mo (kb_orbit.hip)
struct OrbitArgs {
  const double* Mo;  // [Mp][Kp], zero-padded: Mo[o][b] = sum_{x in orbit b} cos(k_o . x)
  const double* H;
  double* Y;         // [Kp][ldg] inputs of (t-1, t') packed per orbit, columns 2q + comp
  double* G;         // [Kp][ldg] pointwise products of (t-1, t'), same layout
  double* S;
  const StepState* st;
  long NP;
  int N, V, Mp, Kp, ldg;  // ldg: 2 (N - 1) rounded up to whole tiles; rows V.. of Y and G are zero
  double preF, preR;
  // 1/N expansion: (Fx, rx) of the bubble pass kept for the chain pass
  // (Y's layout; not kept when NULL), the bubble triangles, the (IF, Ir) ring
  double* X;
  double* PF;
  double* PR;
  const double* Iring;
  long PNP;
  double ivol;
  int classical;  // the sunset without its rho^2 terms (LambdaNLOclassical)
};
// what the two products compute: the sunset (Sigma into the ring), the 1/N
// bubbles of row t-1 (into the bubble triangles), the 1/N chain (Sigma)
enum OrbitStage { ORBIT_SUNSET = 0, ORBIT_BUBBLE = 1, ORBIT_CHAIN = 2 };
constexpr int ORBIT_TILE = 64;  // output tile of the orbit products (rows of Mo padded to it)
constexpr int ORBIT_K = 64;     // k step (columns of Mo padded to it)
size_t orbit_lds();
hipError_t orbit_prepare();  // LDS attribute of the product kernels
hipError_t launch_orbit(const OrbitArgs& a, OrbitStage stage, hipStream_t st);

constexpr int MEMINT_THREADS = 256;  // default workgroup of the memory-integral kernels
constexpr int MEMINT_FINISH_THREADS = 1024;
constexpr int MEMINT_ROWS = 64;  // rows per work item of the adjoint kernels
constexpr int MEAS_THREADS = 256;

}  // namespace qftkb

#pragma GCC visibility pop
