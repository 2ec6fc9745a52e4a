/*
 * qftkb.h -- C ABI of libqftkb.so: the MI355X (gfx950) back-end for the
 * 2PI Kadanoff-Baym time step of the reference's src/TwoPIScalar model
 * (CPUfull storage semantics; LambdaLO, LambdaNLO quantum and classical,
 * NinverseLO and NinverseNLO expansions).
 *
 * The reference selects its numerics back-end by Julia dispatch on the
 * "Num" parameter (TwoPIScalar_Modelfile.jl:104-108) and then calls
 * initSimData!, evolve!, getHartreeMass2 and createMeasurement on the
 * back-end's SimData type.  Each entry point below replaces one of those
 * methods for a new numerics type; the Julia binding a maintainer adds is
 * qftdynamicsm.jl_amd/julia/TwoPIScalar_GPUfull.jl.
 *
 * Conventions
 *   - Every call returns an int status: KB_OK (0) or a KB_ERR_* code; the
 *     message is available from kb_last_error().  No C++ exception crosses
 *     the ABI.
 *   - Times are 1-based as in the reference.  The memory window is
 *     [tmin, tmax] (simdata.indices, src/DefSimDataElements.jl:141-165);
 *     an index outside it returns KB_ERR_WINDOW (the reference throws
 *     BoundsError, DefSimDataElements.jl:187-188).
 *   - Host buffers are caller-owned, contiguous float64 lattices of
 *     V = Nx^sdim sites in column-major site order (x fastest), i.e. the
 *     memory of a Julia Array{Float64,sdim}.
 *   - rho is always exchanged as the physical (antisymmetric) spectral
 *     function rho(i,j); the reference's sign-folded storage
 *     (r[i,j] * thesign(i,j), DefSimDataElements.jl:224-230) is internal.
 *   - The library owns all device memory.  A context is not re-entrant: one
 *     HIP stream, one caller thread.  kb_step enqueues asynchronously;
 *     calls that return data to the host synchronise the stream.
 */
#ifndef QFTKB_H
#define QFTKB_H

#ifdef __cplusplus
extern "C" {
#endif

#define KB_OK 0
#define KB_ERR_ARG 1       /* invalid argument / unsupported size          */
#define KB_ERR_WINDOW 2    /* time index outside the memory window         */
#define KB_ERR_OOM 3       /* device or host allocation failed             */
#define KB_ERR_HIP 4       /* HIP runtime error                            */
#define KB_ERR_STATE 5     /* call out of order (e.g. step before init)    */
#define KB_ERR_NONFINITE 6 /* NaN/Inf detected by kb_check_finite          */

#define KB_PEXP_LAMBDA_LO 0  /* TwoPIScalarLambdaLO (Problemdefinition.jl:34) */
#define KB_PEXP_LAMBDA_NLO 1 /* TwoPIScalarLambdaNLOquantum (:36)            */
#define KB_PEXP_NINVERSE_LO 2  /* TwoPIScalarNinverseLO (:44)             */
#define KB_PEXP_NINVERSE_NLO 3 /* TwoPIScalarNinverseNLO (:45)            */
/* TwoPIScalarLambdaNLOclassical (:39): the sunset without its rho^2 terms,
 * Sigr ~ C[rx Fx^2], SigF ~ C[Fx^3] (TwoPIScalar_Evolution.jl:107-120; the
 * reference defines it for its reduced CPUred storage, here it runs on full
 * and cubic-orbit storage alike) */
#define KB_PEXP_LAMBDA_NLO_CLASSICAL 4

typedef struct kb_ctx kb_ctx;

/* Problem description.  Mirrors TwoPIScalarDiscretizationLattice
 * (TwoPIScalar_Discretization.jl:4-17) and Phi4 (ModelExpansion-
 * Renormalization.jl:5-9).  Mass2 is the model's (possibly renormalised)
 * Mass2; Mass is the discretisation mass used for omega^2 at set-up and for
 * the measurement time. */
typedef struct kb_params {
  int Nx;             /* sites per dimension, 1 .. 2048; Nx^sdim <= 2^24 */
  int sdim;           /* 1, 2 or 3                                       */
  int NstepsinMemory; /* memory window length N_mem, 3 .. 640            */
  int ONgroup;        /* N of O(N)                                       */
  double dt;
  double Lambda;
  double Mass;
  double Mass2;
  int pexp;   /* KB_PEXP_*                                                */
  int device; /* HIP device ordinal                                       */
} kb_params;

/* Scalar part of MeasurementTwoPIScalar (TwoPIScalar_Measurement.jl:2-23). */
typedef struct kb_meas {
  double time, E, E_pot, E_kin, E_mem, M2;
  int n_unstable; /* sites where 1 - dt^2 omega^2/4 < 0 (the reference
                     prints "E" and clamps c = 1, Measurement.jl:48-53) */
} kb_meas;

/* Replaces getTwoPISimData / getTwoPIScalarTmpData for the new numerics
 * type (TwoPIScalar_SimDataStruct.jl:669-721): allocates the device-resident
 * two-time window of F and rho, the self-energy ring and scratch. */
int kb_create(const kb_params *p, kb_ctx **out);
int kb_destroy(kb_ctx *ctx);
const char *kb_last_error(const kb_ctx *ctx);

/* Number of lattice sites V and bytes of device memory held.  (kb_last_error
 * with a NULL context returns the calling thread's last message of a
 * context-free call, kb_sunset below.) */
int kb_info(const kb_ctx *ctx, long long *V, long long *device_bytes);

/* Cubic-orbit storage: the same KB step with one stored value per orbit of
 * the lattice's cubic group (axis permutations and reflections
 * n_i -> Nx - n_i), about 48x fewer stored values in 3D (969 orbits for
 * the 32768 sites of 32^3).  Exact for cubic-symmetric states (the step
 * maps them to cubic-symmetric states).  The self-energy transforms run as
 * FP64 matrix products with the orbit cosine matrix (up to 4096 orbits,
 * KB_FFT_ORBIT_PRODUCT below) or as plane FFTs over the whole lattice
 * (Nx <= 64; in 1D up to 2048).  Every entry point keeps exchanging whole
 * lattices of V sites: inputs are read at each orbit's representative site
 * (the caller supplies symmetric data), outputs are expanded over the
 * orbits.  KB_ERR_ARG for a transform stage with neither form available
 * (Nx > 64 in 2D/3D and more than 4096 orbits); no adjoint.  Ensembles take such
 * members through kb_ens_create_sym.  The reference's CPUred
 * storage (TwoPIScalar_SimDataStruct.jl:204-260) groups sites by FFTWhelper
 * levels instead; this is the exact-symmetry version of that idea. */
int kb_create_sym(const kb_params *p, kb_ctx **out);
/* Stored sites (orbits, or V) and lattice sites V. */
int kb_storage_info(const kb_ctx *ctx, long long *stored_sites,
                    long long *lattice_sites);

/* Per-site lattice momentum squared kL2 (SimDataStruct.jl:673-679).  The
 * library fills it at creation with sum_i 4 sin^2(pi n_i / Nx); a caller
 * holding FFTWhelper levels may overwrite it. */
int kb_get_kL2(kb_ctx *ctx, double *kL2);
int kb_set_kL2(kb_ctx *ctx, const double *kL2);

/* initSimData! (Initialization.jl:8-29): from per-site occupations n[V] and
 * the initial mass, write F(1,1), F(2,1), F(2,2), rho(2,1) = dt, then the
 * Hartree mass at (2,2).  The window becomes [1,2]. */
int kb_init_particles(kb_ctx *ctx, const double *n, double initmass);
/* initSimData! with the three F lattices supplied (Gauss init,
 * Initialization.jl:95-112, or any raw Gaussian state). */
int kb_init_rows(kb_ctx *ctx, const double *F11, const double *F21,
                 const double *F22);

/* evolve! (TwoPIScalar_Evolution.jl:924-981): one KB step producing row t
 * of F and rho.  t must equal tmax + 1 and be >= 3. */
int kb_step(kb_ctx *ctx, int t);

/* Replay each step from a captured hipGraph (on = 1): the first kb_step
 * after enabling captures the step's kernels (advance, self-energy,
 * Hartree, memory integral), later steps launch the graph.  Ignored while
 * timing is enabled.  No counterpart in the reference (launch mechanics). */
int kb_use_graph(kb_ctx *ctx, int on);

/* simdata.indices after the last call. */
int kb_window(const kb_ctx *ctx, int *tmin, int *tmax);

/* simdata.hartreemass2[1]: the Hartree mass^2 used by the last step (or set
 * by init).  Synchronises. */
int kb_hartree(kb_ctx *ctx, double *m2);

/* createMeasurement (TwoPIScalar_Measurement.jl:31-99) at time t (t must be
 * tmax).  Energies are exact sums over all V sites.  Per-site outputs may be
 * NULL; otherwise each receives V doubles:
 *   F = F(t-1,t-1), ddF, r = rho(t-1,t-1), n, omega, c.  Synchronises. */
int kb_measure(kb_ctx *ctx, int t, kb_meas *out, double *F, double *ddF,
               double *r, double *n, double *omega, double *c);
/* The history part of the same record (Measurement.jl:90-96), at site 1:
 * SigF, Sigr for tp = t-1 down to tmin, Fonet = F(tmin,tp) and
 * ronet = rho(tmin,tp) for tp = tmin .. t-1.  Each buffer holds
 * t - tmin doubles; any may be NULL. */
int kb_measure_history(kb_ctx *ctx, int t, double *SigF, double *Sigr,
                       double *Fonet, double *ronet);

/* Test / checkpoint access to the window (which: 0 = F, 1 = rho). */
int kb_get_entry(kb_ctx *ctx, int which, int i, int j, double *out);
int kb_set_entry(kb_ctx *ctx, int which, int i, int j, const double *in);
/* Self-energy lattices of the last step: which 0 = SigF, 1 = Sigr, at tp. */
int kb_get_sigma(kb_ctx *ctx, int which, int tp, double *out);
/* NinverseNLO internals, in momentum space (the CPUred2 quantities of the
 * reference): which = 0 the bubble F2kr2(i,j), 1 the bubble Fr(i,j)
 * (physical, antisymmetric), 2 the resummed kernel Ir(i,j) for i > j
 * (calc_Fr_F2kr2!, SimDataStruct.jl:726-750; Irk, Evolution.jl:252-292);
 * 3 IF(tmax-1, j) and 4 Ir(tmax-1, j) of the last step (Irk/IFk,
 * Evolution.jl:252-292, 539-597; the reference keeps their transforms
 * Irx/IFx, SimDataStruct.jl:70-71), i ignored.  KB_ERR_STATE for other
 * expansions. */
int kb_ninv_get(kb_ctx *ctx, int which, int i, int j, double *out);
/* Benchmark support, no reference counterpart: the streaming-read rate of
 * device `device` (GB/s) measured with non-temporal 16-byte loads over a
 * buffer of `bytes` (at least 16 MiB), the practical ceiling of the
 * bandwidth-bound kernels.  Allocates and frees the buffer. */
int kb_read_bandwidth(int device, long long bytes, double *gbs);
/* Scan the window for NaN/Inf; returns KB_ERR_NONFINITE if any. */
int kb_check_finite(kb_ctx *ctx);

/* Benchmark set-up in the spirit of the reference's benchmark harness
 * (TwoPIScalar_Bechmark.jl:62-69, which fills the window without physics):
 * writes a full, deterministic window [tmax-N_mem+1, tmax] holding the free
 * propagators F = (n+1/2)/w cos(w(a-b)dt), rho = sin(w(a-b)dt)/w with
 * w = sqrt(kL2 + m^2), and sets the window indices.  n[V] may be NULL (0). */
int kb_fill_free_window(kb_ctx *ctx, int tmax, const double *n, double mass);

/* Checkpoint and restart of a run, no counterpart in the reference (its
 * CPUfull SimData lives in process memory): a snapshot of the window, the
 * self-energy ring, kL2, the Hartree mass and the window indices, in the
 * device layout behind a versioned header.  kb_window_bytes gives its size
 * (about the device memory of the window: 42.4 GB at 32^3, N_mem = 400);
 * kb_save_window writes it to a caller buffer of that size (synchronising);
 * kb_load_window restores it into a context created with the same
 * parameters (Nx, sdim, N_mem, N, dt, lambda, masses, expansion, storage;
 * the device may differ), after which kb_step continues the run where the
 * snapshot was taken.  KB_ERR_ARG for a buffer of the wrong size, another
 * format or other parameters; KB_ERR_STATE while recording for the adjoint
 * or for ensemble members.  NinverseNLO rebuilds its bubble and Ir rows at
 * the next step. */
int kb_window_bytes(const kb_ctx *ctx, unsigned long long *bytes);
int kb_save_window(kb_ctx *ctx, void *buf, unsigned long long bytes);
int kb_load_window(kb_ctx *ctx, const void *buf, unsigned long long bytes);

/* Stream and timing. */
int kb_synchronize(kb_ctx *ctx);
/* Launch on a caller-provided hipStream_t (NULL = library's own stream). */
int kb_set_stream(kb_ctx *ctx, void *hip_stream);
/* HIP-event timing of the kernels of kb_step, recorded on the launch
 * stream.  stage: 0 = memory integral + leapfrog kernel, 1 = self-energy
 * stage (all its FFT passes; for NinverseNLO also the resummation), 2 =
 * Hartree reduction, 3 = adjoint memory integral, 4 = the other kernels of
 * the adjoint sweep, 5 = the NinverseNLO resummation sweeps alone.  kb_timing
 * synchronises and returns the accumulated milliseconds and launch count
 * since the last reset. */
int kb_enable_timing(kb_ctx *ctx, int on);
int kb_timing(kb_ctx *ctx, int stage, double *total_ms, long long *count);
int kb_reset_timing(kb_ctx *ctx);

/* ---- tile autotuning -------------------------------------------------------
 * The memory-integral kernel exists in several tile shapes (slot rows per
 * work item, rows per batch, threads per workgroup).  kb_autotune re-times the last step's memory
 * integral in every shape (reps launches each, HIP events on the context's
 * stream; each launch recomputes row t from the same inputs, so row t may
 * change in the last bits) and keeps the fastest; ms, if not NULL, receives
 * one average per shape.  Shapes can also be set from a stored choice.  No
 * counterpart in the reference (launch configuration). */
/* rows per work item, rows per batch, threads per workgroup of a shape */
int kb_tile_shape(int variant, int *rows, int *batch, int *threads); /* KB_ERR_ARG past the last */
int kb_set_tiles(kb_ctx *ctx, int variant);
int kb_get_tiles(const kb_ctx *ctx, int *variant);
int kb_autotune(kb_ctx *ctx, int reps, int *best, double *ms);
/* The same for the self-energy stage's plane kernels (threads per
 * workgroup); re-running the stage recomputes the same self-energies.
 * Orbit-storage LambdaNLO and NinverseNLO contexts (kb_create_sym, up to
 * 4096 orbits) have one more form, KB_FFT_ORBIT_PRODUCT, their default: the
 * transforms of the sunset (or of the 1/N bubble and chain passes) as FP64
 * matrix products with the orbit cosine matrix.  In LambdaNLO contexts
 * kb_autotune_fft times it as well, writes its average after the plane
 * shapes' (ms needs kb_fft_shape's count + 1 entries) and *best is
 * KB_FFT_ORBIT_PRODUCT when it is the fastest. */
#define KB_FFT_ORBIT_PRODUCT (-1)
int kb_fft_shape(int variant, int *threads); /* KB_ERR_ARG past the last */
int kb_set_fft_tiles(kb_ctx *ctx, int variant);
int kb_get_fft_tiles(const kb_ctx *ctx, int *variant);
int kb_autotune_fft(kb_ctx *ctx, int reps, int *best, double *ms);

/* ---- adjoint ---------------------------------------------------------------
 * Gradients of L = sum_p wgt(p) n(p, T) (n as in kb_measure at T = tmax,
 * TwoPIScalar_Measurement.jl:38-59) with respect to the coupling Lambda and
 * the initial state, by hand-written reverse kernels of the whole KB step
 * (evolve!, TwoPIScalar_Evolution.jl:924-981).  The reference has no
 * derivative code; these entry points are what a ChainRulesCore rrule of
 * the back-end's evolve loop calls.  Mass2 is held fixed. */

/* Start recording right after kb_init_particles / kb_init_rows (window
 * [1,2]).  Later kb_step calls run eagerly and keep what the reverse sweep
 * needs for up to max_steps:
 *   checkpoint_every = 0: every column that a step drops from the window
 *                         (N_mem slot pairs, V sites) is kept on a tape;
 *   checkpoint_every = C: the window is kept after steps 2, 2+C, 2+2C, ...
 *                         and each C-step segment is recomputed (taping its
 *                         columns) before it is reversed;
 *   checkpoint_every = -1: the tape of every dropped column, packed to one
 *                         value per orbit of the lattice's cubic group
 *                         (axis permutations, reflections n -> Nx - n).  The
 *                         dynamics keep a cubic-symmetric state symmetric,
 *                         so a site's column is its orbit's; kL2 and the
 *                         initial rows must be orbit-constant to 1e-12
 *                         (KB_ERR_ARG otherwise).  At 48^3 this shrinks the
 *                         tape 38x.
 * Fails with KB_ERR_OOM if the tape or checkpoints do not fit. */
int kb_adj_record(kb_ctx *ctx, int checkpoint_every, int max_steps);
/* kb_adj_record with C = checkpoint_every > 0 for windows that do not fit
 * in device memory C at a time: up to host_windows windows are kept in
 * pinned host memory (each as large as the device history), the initial
 * window as its three rows on the device, and the device tape holds one
 * C-step segment.  The reverse sweep restores the nearest kept window below
 * the segment it needs and recomputes forward from it, keeping new windows
 * on the way by the binomial checkpointing schedule (Griewank's revolve
 * over segments): the least recomputation for host_windows snapshots.
 * host_windows = 0 recomputes every segment from the initial window.  Same
 * gradient as the other modes; kb_adj_info counts the recomputed steps.
 * No counterpart in the reference (it has no derivative code).  KB_ERR_OOM
 * when the pinned host memory cannot be allocated. */
int kb_adj_record_host(kb_ctx *ctx, int checkpoint_every, int max_steps, int host_windows);
/* Reverse sweep from T = tmax down to the initial rows.  wgt[V] are the
 * loss weights.  Outputs (any may be NULL): dL/dLambda; dL/dn(p,0) per site
 * (needs kb_init_particles; n enters F(1,1) = F(2,1) = (n+1/2)/w and
 * F(2,2) = (n+1/2)(w dt^2 + 1/w), Initialization.jl:8-29); dL/dF of the
 * initial rows as 3 x V doubles (F(1,1), F(2,1), F(2,2)).  Afterwards the
 * window is the initial state [1,2] again and recording continues. */
int kb_adj_gradient(kb_ctx *ctx, const double *wgt, double *dL_dlambda,
                    double *dL_dn0, double *dL_dF0);
/* Device bytes held for the adjoint and forward steps recomputed by the
 * last kb_adj_gradient. */
int kb_adj_info(kb_ctx *ctx, long long *device_bytes,
                long long *recomputed_steps);
/* Stop recording and free the adjoint state. */
int kb_adj_stop(kb_ctx *ctx);
/* Test access.  kb_adj_debug_stop makes the next kb_adj_gradient return
 * right after reversing step t_stop (window of step t_stop - 1, no outputs
 * written); kb_adj_get_entry then reads the adjoint of F(i,j) (which = 0)
 * or of rho(i,j) (which = 1) accumulated so far.  The context must be
 * re-initialised afterwards. */
int kb_adj_debug_stop(kb_ctx *ctx, int t_stop);
int kb_adj_get_entry(kb_ctx *ctx, int which, int i, int j, double *out);

/* ---- self-energies of caller lattices ----------------------------------------
 * calcSelfEnergyattp! (TwoPIScalar_Evolution.jl:25-31, LambdaNLOquantum) on
 * lattices that the caller holds in device memory, without a context: for
 * each of nrows rows q, Sigr(q) and SigF(q) of the pair F(q), rho(q) (rho
 * physical), by the plane kernels of the KB step with the same twiddles, so
 * a row gives the same bits as kb_step's self-energy stage.  Rows are V =
 * Nx^sdim doubles at q*V, column-major sites; F, rho, SigF and Sigr are
 * device pointers on p->device.  Only Nx, sdim, ONgroup, Lambda, pexp
 * (KB_PEXP_LAMBDA_NLO, or KB_PEXP_LAMBDA_NLO_CLASSICAL for the classical
 * sunset) and device of *p are read; Nx <= 64, 1 <= nrows <= 65535.  The
 * outputs may not overlap the inputs or each other.
 *
 * The workspace is caller-owned device memory of kb_sunset_workspace_bytes
 * (4 KiB of twiddles, plus nrows * Nx^2 (Nx/2 + 1) complex values in 3D: the
 * half space of each row's transforms), 256-byte
 * aligned.  Kernels are enqueued on `stream` (a hipStream_t; NULL = the
 * device's null stream) and the call returns without waiting.  Errors are
 * reported by the return code; kb_last_error(NULL) gives the calling
 * thread's last message of a context-free call.  Calls that share a
 * workspace must not overlap. */
int kb_sunset_workspace_bytes(const kb_params *p, int nrows, unsigned long long *bytes);
int kb_sunset(const kb_params *p, int nrows, const double *F, const double *rho, double *SigF,
              double *Sigr, void *workspace, unsigned long long workspace_bytes, void *stream);
/* kb_sunset with no workspace argument and no allocation: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It makes no hipMalloc/hipFree or other allocating call.  The workspace is
 * a 64 MiB device array inside the library, one per device, held until the
 * process ends: KB_ERR_ARG when kb_sunset_workspace_bytes of the request
 * exceeds 64 MiB (more than 31 rows of 64^3, 72 of 48^3, 240 of 32^3; 1D and
 * 2D lattices need only the 4 KiB of twiddles).
 * Every call on one device shares that array, so calls that overlap in time
 * on one device (other streams, other threads) race; the results equal
 * kb_sunset's. */
int kb_sunset_noalloc(const kb_params *p, int nrows, const double *F, const double *rho,
                      double *SigF, double *Sigr, void *stream);

/* ---- ensembles ------------------------------------------------------------
 * Many independent KB runs (uneven lattices, couplings, memory windows)
 * advanced together: every member keeps its own device state; one ensemble
 * step launches each member's self-energy and Hartree kernels (on up to four
 * streams at once) and ONE memory-integral launch over the site blocks of
 * all members (an ensemble whose members are all in orbit storage launches
 * it per stream instead, right after that stream's stages).  The
 * reference runs such ensembles as separate jobs (G_SimulationArgs.jl,
 * submit*.sh); there is no ensemble type to mirror. */
typedef struct kb_ens kb_ens;

/* nruns parameter sets; all members live on params[0].device. */
int kb_ens_create(const kb_params *params, int nruns, kb_ens **out);
/* The same with a storage per member: symmetric[i] != 0 creates member i in
 * cubic-orbit storage (as kb_create_sym), 0 stores every lattice site;
 * symmetric NULL: every member in orbit storage.  The memory-integral launch
 * covers each member's stored sites; a member's self-energy stage is the one
 * kb_create_sym picks for it. */
int kb_ens_create_sym(const kb_params *params, const int *symmetric, int nruns, kb_ens **out);
int kb_ens_destroy(kb_ens *ens);
const char *kb_ens_last_error(const kb_ens *ens);
/* Member i as a kb_ctx for initialisation, measurement and read-out
 * (kb_init_*, kb_measure, kb_get_entry, ...).  kb_step on a member is
 * rejected; members advance only through kb_ens_step. */
int kb_ens_member(kb_ens *ens, int i, kb_ctx **member);
/* Every member takes step t (all members must be at tmax = t - 1). */
int kb_ens_step(kb_ens *ens, int t);
/* Capture one ensemble step as a hipGraph and replay it (on = 1). */
int kb_ens_use_graph(kb_ens *ens, int on);
int kb_ens_synchronize(kb_ens *ens);
/* HIP-event time of whole ensemble steps since the last reset. */
int kb_ens_enable_timing(kb_ens *ens, int on);
int kb_ens_timing(kb_ens *ens, double *total_ms, long long *count);
/* Shape of the pooled memory-integral launch (a kb_tile_shape variant; 0
 * at creation), one for all members.  kb_ens_autotune re-times the last
 * ensemble step's memory integral in every shape (reps launches each; each
 * recomputes row t of every member from the same inputs, so row t may change
 * in the last bits) and keeps the fastest; ms, if not NULL, receives one
 * average per shape.  KB_ERR_STATE before the first kb_ens_step or after a
 * member was re-initialised.  No counterpart in the reference (launch
 * configuration). */
int kb_ens_set_tiles(kb_ens *ens, int variant);
int kb_ens_get_tiles(const kb_ens *ens, int *variant);
int kb_ens_autotune(kb_ens *ens, int reps, int *best, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* QFTKB_H */
