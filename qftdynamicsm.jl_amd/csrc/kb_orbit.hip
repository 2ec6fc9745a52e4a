// kb_orbit.hip -- the sunset self-energy of a cubic-orbit-stored window as two
// FP64 matrix products on the matrix cores.
//
// In orbit storage every lattice function is constant on the orbits of the
// cubic group (axis permutations, reflections n -> Nx - n), and so is its
// transform.  The unnormalised cosine transform C X(k) = sum_x X(x) cos(k.x)
// that the sunset uses (TwoPIScalar_Evolution.jl:25-31, real(rfft) of an even
// lattice) restricted to orbit-constant functions is a dense matrix on the
// orbits:
//   (C X)(o) = sum_b Mo[o][b] X(b),   Mo[o][b] = sum_{x in orbit b} cos(k_o . x)
// with k_o the representative site of orbit o.  The self-energy stage
//   Fx = C F(t-1, t'),  rx = C rho(t-1, t')            (all live t')
//   Sigma_F = preF C [Fx (Fx^2 - 3/4 rx^2)],  Sigma_r = preR C [rx (Fx^2 - rx^2/12)]
// is then two products Mo . Y with Y the orbits x (F, rho) columns of every
// live t': 2 * 969^2 * 798 flops per product at 32^3, N_mem = 400.
//
// orbit_gather packs row t-1 of the history (F and the sign-folded rho of the
// pairs (t-1, t')) into Y[orbit][2 q + comp]: the pairs of one row lie spread
// over the whole triangle, so they are read once here rather than by every
// tile of the product.  orbit_gemm<MODE> computes one 64 x 64 tile of Mo . Y
// with v_mfma_f64_16x16x4f64 (8 waves, 32 x 16 per wave), staging 64 x 64
// tiles of Mo and of Y in double-buffered LDS; the next k step's tiles are
// loaded into registers while the current one is multiplied.  The first
// product multiplies Y, and its epilogue pairs the two columns of each t'
// across neighbouring lanes and stores the pointwise products G in Y's
// layout; the second multiplies G into the stage's destination.
//
// The 1/N expansion's two transform passes (TwoPIScalar_Evolution.jl:
// 144-735, the bubble and chain passes of the plane kernels) take the same
// form: the bubble pass stores (Fx^2 - rx^2/4, Fx rx) and keeps (Fx, rx) in
// X, then writes row t-1 of the bubble triangles; the chain pass gathers
// (IF, Ir) from its ring, forms (Fx IFx - rx Irx/4, Fx Irx + rx IFx) with
// (IFx, Irx) = C (IF, Ir) / V, and writes Sigma.
//
// MFMA operand layout (checked on gfx950): lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16]; accumulator value v of lane l is D[4 v + l / 16][l % 16].
#include "kb_kernels.h"

namespace qftkb {

namespace {

typedef double od4 __attribute__((ext_vector_type(4)));
typedef double od2 __attribute__((ext_vector_type(2)));  // a first-class value (HIP's double2 is a struct)

constexpr int OT = ORBIT_TILE;  // output tile (rows and columns)
constexpr int OK = ORBIT_K;     // k per LDS stage
constexpr int OTHREADS = 512;
static_assert((OT * OK / 2) % OTHREADS == 0 && OT == 64 && (OK % 4) == 0, "orbit product tiling");

// LDS: two buffers of [A: OT rows x ALD] [B: OK rows x BLD] doubles.  Row
// strides of 66 and 80 doubles keep the lanes of each half-wave on distinct
// banks when the MFMA operands are read, and rows 16-byte aligned for the
// paired stores.
constexpr int ALD = OK + 2, BLD = OT + 16;
constexpr int ASZ = OT * ALD, BSZ = OK * BLD, BUFSZ = ASZ + BSZ;
constexpr int NA = (OT * OK / 2) / OTHREADS;  // double2 of Mo per thread and k step
constexpr int NB = (OK * OT / 2) / OTHREADS;  // double2 of Y per thread and k step

// gather tile: 64 orbits x 32 t'
constexpr int GB = 64, GQ = 32;

}  // namespace

// blockIdx.x: 64 orbits, blockIdx.y: 32 t'.  RING: (IF, Ir)(t-1, t') from
// the 1/N ring (self-energy layout) instead of (F, rho) from the history.
template <bool RING>
__global__ __launch_bounds__(OTHREADS) void orbit_gather(OrbitArgs a) {
  __shared__ double2 t[GQ][GB + 1];
  const StepState ss = *a.st;
  const int b0 = blockIdx.x * GB, q0 = blockIdx.y * GQ;
  if (q0 >= ss.L) return;
  const int tid = threadIdx.x;
  // neighbouring threads read neighbouring orbits of one pair: 16 bytes
  // each, 8 orbits per 128-byte line
#pragma unroll
  for (int e = 0; e < (GB * GQ) / OTHREADS; ++e) {
    const int idx = tid + e * OTHREADS;
    const int bb = idx % GB, qq = idx / GB;
    const int b = b0 + bb, q = q0 + qq;
    double2 v = make_double2(0.0, 0.0);
    if (b < a.V && q < ss.L) {
      const int slot = (ss.tmin + q - 1) % a.N;
      if (RING) {
        v = *reinterpret_cast<const double2*>(a.Iring + soff(b / MB, slot, b % MB, a.N));
      } else {
        v = *reinterpret_cast<const double2*>(a.H + hoff(b / MB, pair_off(ss.s_tm1, slot, a.N), b % MB, a.NP));
        if (ss.s_tm1 < slot) v.y = -v.y;
      }
    }
    t[qq][bb] = v;
  }
  __syncthreads();
  // neighbouring threads write neighbouring t' of one orbit
#pragma unroll
  for (int e = 0; e < (GB * GQ) / OTHREADS; ++e) {
    const int idx = tid + e * OTHREADS;
    const int qq = idx % GQ, bb = idx / GQ;
    const int b = b0 + bb, q = q0 + qq;
    if (b < a.V && q < ss.L) *reinterpret_cast<double2*>(a.Y + (size_t)b * a.ldg + 2 * q) = t[qq][bb];
  }
}

// MODE: first products (from Y) of the sunset, bubble and chain passes,
// then the second products (from G) into the ring or the bubble triangles
enum OrbitMode { OM_SUNSET, OM_BUBBLE, OM_CHAIN, OM_TO_RING, OM_TO_BUBBLES };

template <int MODE>
__global__ __launch_bounds__(OTHREADS) void orbit_gemm(OrbitArgs a) {
  extern __shared__ double olds[];
  const StepState ss = *a.st;
  const int L = ss.L;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int row0 = blockIdx.x * OT;  // orbits (output rows)
  const int col0 = blockIdx.y * OT;  // columns 2 q + comp
  if (col0 >= 2 * L) return;         // whole tile past the live rows
  constexpr bool FIRST = MODE == OM_SUNSET || MODE == OM_BUBBLE || MODE == OM_CHAIN;
  const double* Bsrc = FIRST ? a.Y : a.G;
  // 8 waves, 2 per SIMD (one multiplies while the other waits for LDS):
  // wave w owns rows wr .. wr + 31 and columns wc .. wc + 15 of the tile
  const int wr = (wave >> 2) * 32, wc = (wave & 3) * 16;
  od4 acc[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) acc[i] = od4{0.0, 0.0, 0.0, 0.0};

  // Tiles of one k step, loaded from global memory into registers one step
  // ahead of the products (double-buffered LDS, one barrier per k step).
  od2 ra[NA];  // Mo: neighbouring columns of one row
  od2 rb[NB];  // Y or G: (F, rho) or (gF, gr) of one orbit and t'
  auto load = [&](int k0) {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      const int idx = tid + e * OTHREADS;
      const int row = idx / (OK / 2), kk = 2 * (idx % (OK / 2));
      ra[e] = *reinterpret_cast<const od2*>(a.Mo + (size_t)(row0 + row) * a.Kp + k0 + kk);
    }
    // neighbouring threads read neighbouring t' of one orbit.  Y and G are
    // padded to Kp rows and whole tiles of columns: rows b >= V stay zero
    // (and meet zero columns of Mo), columns q >= L feed only output columns
    // that are not stored, so the loads need no mask
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      const int idx = tid + e * OTHREADS;
      const int k = idx / (OT / 2), qq = idx % (OT / 2);
      rb[e] = *reinterpret_cast<const od2*>(Bsrc + (size_t)(k0 + k) * a.ldg + col0 + 2 * qq);
    }
  };
  // LDS is indexed from the kernel's own __shared__ array so that every
  // access stays an LDS instruction (a pointer picked at run time would turn
  // them into flat accesses, whose waits also wait for the prefetch)
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      const int idx = tid + e * OTHREADS;
      const int row = idx / (OK / 2), kk = 2 * (idx % (OK / 2));
      *reinterpret_cast<od2*>(olds + buf * BUFSZ + row * ALD + kk) = ra[e];
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      const int idx = tid + e * OTHREADS;
      const int k = idx / (OT / 2), qq = idx % (OT / 2);
      *reinterpret_cast<od2*>(olds + buf * BUFSZ + ASZ + k * BLD + 2 * qq) = rb[e];
    }
  };
  load(0);
  store(0);
  __syncthreads();
  const int nk = a.Kp / OK;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * OK);
    const int abase = cur * BUFSZ + (wr + (lane & 15)) * ALD + (lane >> 4);
    const int bbase = cur * BUFSZ + ASZ + (lane >> 4) * BLD + wc + (lane & 15);
#pragma unroll
    for (int ks = 0; ks < OK; ks += 4) {
      double av[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) av[i] = olds[abase + 16 * i * ALD + ks];
      const double bv = olds[bbase + ks * BLD];
#pragma unroll
      for (int i = 0; i < 2; ++i) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv, acc[i], 0, 0, 0);
    }
    // the other buffer was last read before the previous barrier
    if (kt + 1 < nk) store(cur ^ 1);
    __syncthreads();
  }

  // epilogue: value v of sub-tile i is row wr + 16 i + 4 v + lane / 16,
  // column wc + lane % 16; columns 2q (F) and 2q+1 (rho) sit in
  // neighbouring lanes
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int o = row0 + wr + 16 * i + 4 * v + (lane >> 4);
      const int col = col0 + wc + (lane & 15);
      const int q = col >> 1, comp = col & 1;
      const double mine = acc[i][v];
      const double other = __shfl_xor(mine, 1);
      const bool live = o < a.V && q < L;
      const size_t gi = (size_t)o * a.ldg + col;
      if constexpr (MODE == OM_SUNSET) {
        // quantum sunset; the classical one drops the rho^2 terms
        const double qF = a.classical ? 0.0 : 0.75, qr = a.classical ? 0.0 : 1.0 / 12.0;
        const double Fx = comp ? other : mine, rx = comp ? mine : other;
        const double FF = Fx * Fx, rr = rx * rx;
        const double g = comp ? rx * (FF - qr * rr) : Fx * (FF - qF * rr);
        if (live) a.G[gi] = g;
      } else if constexpr (MODE == OM_BUBBLE) {
        // F2kr2 and Fr integrands: Fx^2 - rx^2 / 4, Fx rx
        const double Fx = comp ? other : mine, rx = comp ? mine : other;
        const double g = comp ? Fx * rx : Fx * Fx - 0.25 * (rx * rx);
        if (live) {
          a.G[gi] = g;
          if (a.X) a.X[gi] = mine;
        }
      } else if constexpr (MODE == OM_CHAIN) {
        // (Fx IFx - rx Irx / 4, Fx Irx + rx IFx) with (IFx, Irx) = C (IF, Ir) / V
        const double IFx = a.ivol * (comp ? other : mine), Irx = a.ivol * (comp ? mine : other);
        double Fx = 0.0, rx = 0.0;
        if (live) {
          const size_t xi = (size_t)o * a.ldg + 2 * q;
          Fx = a.X[xi];
          rx = a.X[xi + 1];
        }
        const double g = comp ? Fx * Irx + rx * IFx : Fx * IFx - 0.25 * (rx * Irx);
        if (live) a.G[gi] = g;
      } else if constexpr (MODE == OM_TO_RING) {
        if (live) {
          const int slot = (ss.tmin + q - 1) % a.N;
          a.S[soff(o / MB, slot, o % MB, a.N) + comp] = (comp ? a.preR : a.preF) * mine;
        }
      } else {  // OM_TO_BUBBLES: row t-1 of the bubble triangles, Fr sign-folded like rho
        if (live) {
          const int slot = (ss.tmin + q - 1) % a.N;
          const size_t pi = poff(o / MB, ppair_off(ss.s_tm1, slot, a.N), o % MB, a.PNP);
          if (comp)
            a.PR[pi] = (ss.s_tm1 >= slot ? a.preR : -a.preR) * mine;
          else
            a.PF[pi] = a.preF * mine;
        }
      }
    }
}

size_t orbit_lds() { return sizeof(double) * 2 * (size_t)BUFSZ; }

hipError_t orbit_prepare() {
  const void* fns[] = {(const void*)orbit_gemm<OM_SUNSET>, (const void*)orbit_gemm<OM_BUBBLE>,
                       (const void*)orbit_gemm<OM_CHAIN>, (const void*)orbit_gemm<OM_TO_RING>,
                       (const void*)orbit_gemm<OM_TO_BUBBLES>};
  hipError_t e = hipSuccess;
  for (const void* f : fns)
    if (e == hipSuccess) e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)orbit_lds());
  return e;
}

// gather of the stage's inputs, the first product with the pointwise step,
// the second product into its destination
hipError_t launch_orbit(const OrbitArgs& a, OrbitStage stage, hipStream_t st) {
  const int rows = a.N - 1;  // live t' (rows q >= L exit at once)
  const dim3 ggrid((a.V + GB - 1) / GB, (rows + GQ - 1) / GQ);
  if (stage == ORBIT_CHAIN)
    hipLaunchKernelGGL(orbit_gather<true>, ggrid, dim3(OTHREADS), 0, st, a);
  else
    hipLaunchKernelGGL(orbit_gather<false>, ggrid, dim3(OTHREADS), 0, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const dim3 grid(a.Mp / OT, (2 * rows + OT - 1) / OT);
  const size_t lds = orbit_lds();
  switch (stage) {
    case ORBIT_SUNSET: hipLaunchKernelGGL(orbit_gemm<OM_SUNSET>, grid, dim3(OTHREADS), lds, st, a); break;
    case ORBIT_BUBBLE: hipLaunchKernelGGL(orbit_gemm<OM_BUBBLE>, grid, dim3(OTHREADS), lds, st, a); break;
    default: hipLaunchKernelGGL(orbit_gemm<OM_CHAIN>, grid, dim3(OTHREADS), lds, st, a); break;
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (stage == ORBIT_BUBBLE)
    hipLaunchKernelGGL(orbit_gemm<OM_TO_BUBBLES>, grid, dim3(OTHREADS), lds, st, a);
  else
    hipLaunchKernelGGL(orbit_gemm<OM_TO_RING>, grid, dim3(OTHREADS), lds, st, a);
  return hipGetLastError();
}

}  // namespace qftkb
