// kb_lattice.hip -- context-free entry points on lattices that the caller
// keeps in device memory: the sunset self-energies of calcSelfEnergyattp!
// (TwoPIScalar_Evolution.jl:25-31) for rows of F and rho, computed by the
// plane kernels of the KB step (kb_sigma.hip, mode SG_PLAIN) with the same
// twiddles as a context, so a row gives the same bits as the step's
// self-energy stage.
//
// Device memory: the caller's input and output rows, plus a workspace that
// holds the twiddle table (the first WS_TWIDDLE_BYTES) and, for 3D lattices,
// half a complex lattice per row (the planes z in [0, Nx/2]) for the
// transforms between the plane passes.
// kb_sunset takes the workspace from the caller; kb_sunset_noalloc uses the
// file-scope device array `scratch` below (SCRATCH_BYTES on every device the
// library runs on, held until the process ends), so it allocates nothing.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/qftkb.h"
#include "kb_internal.h"

using namespace qftkb;
using namespace qftkb_host;

#pragma GCC visibility push(hidden)
namespace {

thread_local std::string g_err = "no error";

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

constexpr int NX_MAX = 64;                 // the plane kernels' largest lattice
constexpr size_t WS_TWIDDLE_BYTES = 4096;  // >= NX_MAX complex values, keeps the rows 256-byte aligned
constexpr int MAX_DEVICES = 64;
constexpr int MAX_ROWS = 65535;            // grid rows of one launch

// kb_sunset_noalloc's workspace.  In host code `scratch` names a host-side
// stand-in, not a device address: only the address that hipGetSymbolAddress
// returns for the current device (kept per device in g_scratch) is used.
// Nothing in it needs clearing: every call writes the twiddles and the rows
// before reading them.  Calls on one device share it and must not overlap.
constexpr size_t SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) char scratch[SCRATCH_BYTES];

// the host twiddle table, passed by value to the kernel that stores it
struct Twiddles {
  double2 w[NX_MAX];
};

__global__ void put_twiddles(Twiddles t, double2* dst, int n) {
  const int k = threadIdx.x;
  if (k < n) dst[k] = t.w[k];
}

// rows of complex values between the plane passes (3D only): the half space
// z in [0, Nx/2] of each row
size_t workspace_bytes(const kb_params* p, int nrows) {
  const size_t half = (size_t)p->Nx * p->Nx * (p->Nx / 2 + 1);
  return WS_TWIDDLE_BYTES + (p->sdim == 3 ? (size_t)nrows * half * sizeof(double2) : 0);
}

int check_params(const kb_params* p, int nrows) {
  if (!p) return fail(KB_ERR_ARG, "sunset: null parameters");
  if (p->Nx < 1 || p->Nx > NX_MAX) return fail(KB_ERR_ARG, "sunset: Nx = %d outside 1 .. %d", p->Nx, NX_MAX);
  if (p->sdim < 1 || p->sdim > 3) return fail(KB_ERR_ARG, "sunset: sdim = %d", p->sdim);
  if (p->ONgroup < 1) return fail(KB_ERR_ARG, "sunset: ONgroup = %d", p->ONgroup);
  if (!sunset_expansion(p->pexp))
    return fail(KB_ERR_ARG, "sunset: pexp = %d has no sunset self-energy", p->pexp);
  if (nrows < 1 || nrows > MAX_ROWS) return fail(KB_ERR_ARG, "sunset: nrows = %d outside 1 .. %d", nrows, MAX_ROWS);
  if (p->device < 0 || p->device >= MAX_DEVICES) return fail(KB_ERR_ARG, "sunset: device %d", p->device);
  return KB_OK;
}

// [a, a + na) and [b, b + nb) share a byte
bool overlap(const void* a, size_t na, const void* b, size_t nb) {
  const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
  return x < y + nb && y < x + na;
}
bool overlap(const void* a, const void* b, size_t bytes) { return overlap(a, bytes, b, bytes); }

int check_buffers(const double* F, const double* rho, const double* SigF, const double* Sigr, size_t bytes) {
  if (!F || !rho || !SigF || !Sigr) return fail(KB_ERR_ARG, "sunset: null lattice pointer");
  if (overlap(SigF, Sigr, bytes) || overlap(SigF, F, bytes) || overlap(SigF, rho, bytes) ||
      overlap(Sigr, F, bytes) || overlap(Sigr, rho, bytes))
    return fail(KB_ERR_ARG, "sunset: output rows overlap other rows");
  return KB_OK;
}

// the current device for the duration of a call
struct DeviceScope {
  int prev = -1;
  hipError_t e = hipSuccess;
  explicit DeviceScope(int dev) {
    e = hipGetDevice(&prev);
    if (e == hipSuccess && prev != dev) e = hipSetDevice(dev);
  }
  ~DeviceScope() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

// once per device, with that device current: the dynamic-LDS limits of the
// plane kernels and the device address of `scratch`
std::mutex g_setup_mu;
bool g_ready[MAX_DEVICES] = {};
char* g_scratch[MAX_DEVICES] = {};

int device_setup(int dev) {
  std::lock_guard<std::mutex> lk(g_setup_mu);
  if (g_ready[dev]) return KB_OK;
  void* ptr = nullptr;
  hipError_t se = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (se != hipSuccess || !ptr)
    return fail(KB_ERR_HIP, "sunset: no device address for the scratch array: %s", hipGetErrorString(se));
  g_scratch[dev] = static_cast<char*>(ptr);
  // every compiled length and the runtime-length instantiation
  static const int lengths[] = {8, 16, 24, 32, 40, 48, 64, 1};
  for (int n : lengths) {
    hipError_t e = sigma_planes_prepare(n);
    if (e != hipSuccess) return fail(KB_ERR_HIP, "sunset: plane kernel set-up failed: %s", hipGetErrorString(e));
  }
  g_ready[dev] = true;
  return KB_OK;
}

// the three plane passes over `nrows` rows; ws holds the twiddles and the rows
int run_sunset(const kb_params* p, int nrows, const double* F, const double* rho, double* SigF, double* Sigr,
               char* ws, hipStream_t st) {
  const int n = p->Nx;
  Twiddles tw{};
  fft_twiddles(n, tw.w);
  double2* dtw = reinterpret_cast<double2*>(ws);
  hipLaunchKernelGGL(put_twiddles, dim3(1), dim3(NX_MAX), 0, st, tw, dtw, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(KB_ERR_HIP, "sunset: twiddle launch failed: %s", hipGetErrorString(e));
  SigmaArgs a{};
  a.W = reinterpret_cast<double2*>(ws + WS_TWIDDLE_BYTES);
  a.tw = dtw;
  plan_fft(a.plan, n);
  a.n = a.nx = n;
  a.ny = p->sdim >= 2 ? n : 1;
  a.nz = p->sdim >= 3 ? n : 1;
  a.V = a.nx * a.ny * a.nz;
  const double ivol = 1.0 / (double)a.V;
  const double lam = p->Lambda, ON = (double)p->ONgroup;
  a.preF = ivol * ivol * ivol * (-(lam * lam) * (ON + 2.0) / (18.0 * ON * ON));
  a.preR = ivol * ivol * ivol * (-(lam * lam) * (ON + 2.0) / (6.0 * ON * ON));
  a.Fin = F;
  a.Rin = rho;
  a.SigFout = SigF;
  a.Sigrout = Sigr;
  a.rows = nrows;
  a.classical = p->pexp == KB_PEXP_LAMBDA_NLO_CLASSICAL;
  e = launch_sigma_planes_plain(a, st);
  if (e != hipSuccess) return fail(KB_ERR_HIP, "sunset: plane kernel launch failed: %s", hipGetErrorString(e));
  return KB_OK;
}

}  // namespace

namespace qftkb_host {
const char* free_call_error() { return g_err.c_str(); }
}  // namespace qftkb_host
#pragma GCC visibility pop

int kb_sunset_workspace_bytes(const kb_params* p, int nrows, unsigned long long* bytes) {
  int rc = check_params(p, nrows);
  if (rc) return rc;
  if (!bytes) return fail(KB_ERR_ARG, "sunset: null output");
  *bytes = workspace_bytes(p, nrows);
  return KB_OK;
}

int kb_sunset(const kb_params* p, int nrows, const double* F, const double* rho, double* SigF, double* Sigr,
              void* workspace, unsigned long long workspace_bytes_given, void* stream) {
  int rc = check_params(p, nrows);
  if (rc) return rc;
  const size_t lat = (size_t)nrows * p->Nx * (p->sdim >= 2 ? p->Nx : 1) * (p->sdim >= 3 ? p->Nx : 1);
  if ((rc = check_buffers(F, rho, SigF, Sigr, lat * sizeof(double)))) return rc;
  const size_t need = workspace_bytes(p, nrows);
  if (!workspace || workspace_bytes_given < need)
    return fail(KB_ERR_ARG, "sunset: workspace of %llu bytes, %zu needed", workspace_bytes_given, need);
  if ((uintptr_t)workspace % 256 != 0) return fail(KB_ERR_ARG, "sunset: workspace not 256-byte aligned");
  const size_t lb = lat * sizeof(double);
  if (overlap(workspace, need, F, lb) || overlap(workspace, need, rho, lb) || overlap(workspace, need, SigF, lb) ||
      overlap(workspace, need, Sigr, lb))
    return fail(KB_ERR_ARG, "sunset: workspace overlaps the lattices");
  DeviceScope ds(p->device);
  if (ds.e != hipSuccess) return fail(KB_ERR_HIP, "sunset: device %d: %s", p->device, hipGetErrorString(ds.e));
  if ((rc = device_setup(p->device))) return rc;
  return run_sunset(p, nrows, F, rho, SigF, Sigr, static_cast<char*>(workspace), static_cast<hipStream_t>(stream));
}

int kb_sunset_noalloc(const kb_params* p, int nrows, const double* F, const double* rho, double* SigF,
                      double* Sigr, void* stream) {
  int rc = check_params(p, nrows);
  if (rc) return rc;
  const size_t lat = (size_t)nrows * p->Nx * (p->sdim >= 2 ? p->Nx : 1) * (p->sdim >= 3 ? p->Nx : 1);
  if ((rc = check_buffers(F, rho, SigF, Sigr, lat * sizeof(double)))) return rc;
  const size_t need = workspace_bytes(p, nrows);
  if (need > SCRATCH_BYTES)
    return fail(KB_ERR_ARG, "sunset_noalloc: %d rows of %d^%d need %zu bytes of scratch, the limit is %zu", nrows,
                p->Nx, p->sdim, need, SCRATCH_BYTES);
  DeviceScope ds(p->device);
  if (ds.e != hipSuccess) return fail(KB_ERR_HIP, "sunset: device %d: %s", p->device, hipGetErrorString(ds.e));
  if ((rc = device_setup(p->device))) return rc;
  char* ws = g_scratch[p->device];
  const size_t lb = lat * sizeof(double);
  if (overlap(ws, need, F, lb) || overlap(ws, need, rho, lb) || overlap(ws, need, SigF, lb) ||
      overlap(ws, need, Sigr, lb))
    return fail(KB_ERR_ARG, "sunset_noalloc: lattices overlap the library's scratch");
  return run_sunset(p, nrows, F, rho, SigF, Sigr, ws, static_cast<hipStream_t>(stream));
}
