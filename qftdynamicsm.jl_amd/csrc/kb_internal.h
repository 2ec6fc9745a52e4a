// kb_internal.h -- host-side context of libqftkb shared by the forward
// (kb_api.hip) and adjoint (kb_adjoint.hip) translation units.  Nothing here
// is part of the C ABI; every symbol is hidden from the shared library's
// dynamic symbol table.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/qftkb.h"
#include "kb_kernels.h"

#pragma GCC visibility push(hidden)

namespace qftkb_host {

struct EventPair {
  hipEvent_t a, b;
};

// stages of kb_timing: 0 memory integral, 1 self-energy, 2 Hartree,
// 3 adjoint memory integral, 4 the other kernels of the adjoint sweep,
// 5 the 1/N resummation sweeps
constexpr int NSTAGE = 6;

struct AdjState;  // kb_adjoint.hip

}  // namespace qftkb_host

struct kb_ctx {
  kb_params p{};
  int V = 0, Vpad = 0, NB = 0, N = 0;  // V: stored sites (lattice sites, or cubic orbits)
  int Vl = 0;                          // lattice sites Nx^sdim
  // cubic-orbit storage: one stored site per orbit of the lattice's cubic
  // group (axis permutations and reflections n -> Nx - n)
  bool sym = false;
  std::vector<int> orbit_of;  // lattice site -> stored site
  std::vector<int> rep_site;  // stored site -> its representative lattice site
  int* smap_dev = nullptr;     // orbit_of on the device
  int* rmap_dev = nullptr;     // lattice site -> stored site if it is the representative, else -1
  double* wgt_dev = nullptr;   // orbit sizes
  std::vector<double> lat_tmp;
  long NP = 0;
  int nx = 1, ny = 1, nz = 1;
  qftkb::StepState st{};  // host mirror of the window (simdata.indices)
  bool initialised = false;
  bool in_ensemble = false;
  double initmass = 0.0;      // set by kb_init_particles
  bool init_particles = false;
  // device memory
  double* H = nullptr;     // history
  double* S = nullptr;     // self-energy ring
  double2* W = nullptr;    // FFT workspace
  double* kL2 = nullptr;
  double* m2 = nullptr;    // [0] current Hartree mass^2
  double* hpart = nullptr; // Hartree partial sums and their counter (hartree_part_doubles)
  double* partial = nullptr;
  double* modes = nullptr; // 6 x V measurement outputs
  double* vbuf = nullptr;  // 4 x V staging
  double2* tw = nullptr;
  int* flag = nullptr;
  int2* items = nullptr;  // memory-integral work items of 64-row chunks (adjoint kernels, ensembles)
  int nitems = 0;
  int2* items_r[3] = {};  // work items per row-chunk choice (MEMINT_ROW_CHOICES)
  int nitems_r[3] = {};
  int tile = 0;           // memory-integral tile shape (kMemintVariants index)
  int msplit = 1;         // workgroups per site block of the memory integral (split form if > 1)
  double* mpart = nullptr;  // their partial accumulators
  int nsplit = 1;           // the same for the 1/N resummation sweeps
  double* npart = nullptr;  // their partial accumulators
  double* ncv = nullptr;    // coefficient vectors passed between the split sweeps
  int fft_tile = 0;       // self-energy plane kernels' block size (kFFTThreads index),
                          // or KB_FFT_ORBIT_PRODUCT
  // orbit storage, LambdaNLO: the orbit cosine matrix and the sunset products
  double* Mo = nullptr;
  double* Yo = nullptr;  // row t-1 packed per orbit
  double* Xo = nullptr;  // 1/N: (Fx, rx) of the bubble pass, kept for the chain pass
  double* G = nullptr;
  int Mp = 0, Kp = 0, ldg = 0;
  // device-resident step state and launch tables (graph-capturable steps)
  qftkb::StepState* st_dev = nullptr;
  qftkb::StepState** stptr_dev = nullptr;  // {st_dev}
  int* N_dev = nullptr;
  qftkb::MemintArgs* margs_dev = nullptr;
  int* wg0_dev = nullptr;  // {0}
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  size_t bytes = 0;
  qftkb::FFTPlan plan{};
  int fftP = 1;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  bool timing = false;
  std::vector<qftkb_host::EventPair> pending[qftkb_host::NSTAGE];
  std::vector<hipEvent_t> freeEv;
  double tsum[qftkb_host::NSTAGE] = {};
  long long tcount[qftkb_host::NSTAGE] = {};
  qftkb_host::AdjState* adj = nullptr;  // recording / adjoint state
  // 1/N expansion (NinverseNLO): bubble triangles F2kr2 and Fr, the
  // resummed-kernel triangle Ir, the (IF, Ir) ring, the bubble pass's
  // transforms (Fx, rx) for the chain pass, and one step state per row for
  // rebuilding the rows of a window that was written from outside a step
  long PNP = 0;            // slot pairs per site block of those triangles (PCB-wide column blocks)
  int2* pitems = nullptr;  // their sweeps' work items: (column block, 64-row chunk)
  int npitems = 0;
  double* PF = nullptr;
  double* PR = nullptr;
  double* IR = nullptr;
  double* Iring = nullptr;
  double2* Xn = nullptr;
  qftkb::StepState* stfill_dev = nullptr;
  bool ninv_valid = false;  // bubble and Ir rows cover the window
  std::string err;
};

#define HIPCHK(ctx, call)                                                                     \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return seterr((ctx), KB_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                                      \
  } while (0)

#define LAUNCHCHK(ctx)                                                                          \
  do {                                                                                          \
    hipError_t e_ = hipGetLastError();                                                          \
    if (e_ != hipSuccess)                                                                       \
      return seterr((ctx), KB_ERR_HIP, "kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                                        \
  } while (0)

namespace qftkb_host {

int seterr(kb_ctx* c, int code, const char* fmt, ...);
// FFT factorisation and the twiddles exp(-2 pi i k / n), k < n (long double
// on the host), shared by contexts and the context-free entry points
void plan_fft(qftkb::FFTPlan& pl, int n);
void fft_twiddles(int n, double2* out);
// message of the calling thread's last failed context-free call
const char* free_call_error();
int upload(kb_ctx* c, double* dst, const double* src, size_t n);
int download(kb_ctx* c, double* dst, const double* src, size_t n);
// host lattices (Nx^sdim sites) to / from stored sites: identity in full
// storage; representatives / orbit expansion in cubic-orbit storage
int lat_upload(kb_ctx* c, double* dst, const double* src);
int lat_download(kb_ctx* c, double* dst, const double* src);
int dmalloc(kb_ctx* c, void** ptr, size_t bytes);
int upload_state(kb_ctx* c);
void drop_graph(kb_ctx* c);
// self-energy stage of the step in c->st_dev into S_out (default: the
// context's ring); X, if set, receives the transforms (Fx, rx)
int launch_sigma(kb_ctx* c, double2* X = nullptr, double* S_out = nullptr);
// the plane kernels' workspace (W, and Xn for the 1/N passes), allocated
// when a plane form is first selected; and whether it is missing
int ensure_plane_workspace(kb_ctx* c);
bool plane_workspace_missing(const kb_ctx* c);
// Hartreeprefactor of the context's expansion (Initialization.jl:245-247)
double hartree_prefactor(const kb_ctx* c);
// expansions with a self-energy (TwoPIScalarNLO, Problemdefinition.jl:48)
// the lambda expansion's sunset, quantum or classical (TwoPIScalarLambdaNLO,
// Problemdefinition.jl:37)
inline bool sunset_expansion(int pexp) {
  return pexp == KB_PEXP_LAMBDA_NLO || pexp == KB_PEXP_LAMBDA_NLO_CLASSICAL;
}
inline bool has_sigma(const kb_ctx* c) {
  return sunset_expansion(c->p.pexp) || c->p.pexp == KB_PEXP_NINVERSE_NLO;
}
// orbits of the lattice's cubic group (axis permutations, reflections
// n -> Nx - n): lattice site -> orbit, and orbit -> its first site
void cubic_orbits(int Nx, int sdim, std::vector<int>& orbit_of, std::vector<int>& rep_site);
// the kernels of one forward step (the window in c->st_dev is advanced first)
int enqueue_step(kb_ctx* c);
size_t memint_lds(int N);
// launch the forward memory-integral kernel of tile shape `tile` on args/wg tables
hipError_t launch_memint(int tile, unsigned grid, size_t lds, hipStream_t s, const qftkb::MemintArgs* args,
                         const int* wg_start, int nruns);

// times one stage of kb_timing while in scope
struct StageTimer {
  kb_ctx* c;
  int stage;
  EventPair ev{nullptr, nullptr};
  StageTimer(kb_ctx* c_, int s);
  ~StageTimer();
};

// adjoint hooks (kb_adjoint.hip)
int adj_before_step(kb_ctx* c);  // after the host window has advanced to step t
int adj_after_step(kb_ctx* c);
void adj_free(kb_ctx* c);

}  // namespace qftkb_host

#pragma GCC visibility pop
