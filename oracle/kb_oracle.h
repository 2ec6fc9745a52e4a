/*
 * kb_oracle.h -- CPU restatement of the reference 2PI Kadanoff-Baym step
 * (src/TwoPIScalar, CPUfull storage, LambdaNLO "quantum" sunset).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against, and the CPU baseline timed beside it.
 * Nothing in the product (libqftkb.so, the qftkb host package) links,
 * loads or calls it.
 *
 * Conventions follow the reference exactly:
 *   - times are 1-based; the window is [tmin, tmax] (simdata.indices,
 *     src/DefSimDataElements.jl:141-165);
 *   - F and r are SymMatrix ring buffers over N = NstepsinMemory slots,
 *     lower triangle in slot space (DefSimDataElements.jl:169-221);
 *   - the stored spectral function r is sign-folded: the physical
 *     rho(i,j) = r[i,j] * thesign(i,j) (DefSimDataElements.jl:224-230);
 *   - SigF/Sigr are CyclVectors over the same slots (:88-149).
 * Accessors below that say "physical" apply the sign fold.
 */
#ifndef KB_ORACLE_H
#define KB_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

/* pexp: 0 = LambdaLO (no self-energy, zero memory integrals:
 *           Evolution.jl:8-9, 763-765, 842-844),
 *       1 = LambdaNLO quantum (Evolution.jl:25-31),
 *       2 = NinverseLO (no self-energy, Hartree prefactor lambda/6:
 *           Evolution.jl:10-11, Initialization.jl:246),
 *       3 = NinverseNLO (resummed bubble chain, Evolution.jl:144-735),
 *       4 = LambdaNLO classical (the sunset without its rho^2 terms,
 *           Evolution.jl:107-120; the reference defines it for CPUred
 *           storage, the same formulas on the full lattice). */
or_ctx *or_create(int Nx, int sdim, int Nmem, double dt, double lambda,
                  int ONgroup, double Mass, double Mass2, int pexp,
                  int nthreads);
void or_destroy(or_ctx *c);
const char *or_last_error(or_ctx *c);

/* kL2 per site (lev2 restated as sum_i 4 sin^2(pi n_i / Nx)), length V. */
int or_get_kL2(or_ctx *c, double *out);
/* Override the k^2 table (for tests that pin the dynamics with other
 * dispersions).  Also resets omega2 = (kL2 + Mass^2) dt^2
 * (SimDataStruct.jl:673-679). */
int or_set_kL2(or_ctx *c, const double *kL2);

/* initSimData! for CPUfull (Initialization.jl:8-29): expand twice, write
 * F(1,1), F(2,1), F(2,2), r(1,1), r(2,1), r(2,2) from per-site
 * occupations n[V] and the initial mass, then hartreemass2 = M^2(2,2). */
int or_init_particles(or_ctx *c, const double *n, double initmass);
/* Same, with the three F lattices given directly (Gauss init,
 * Initialization.jl:95-112, or any raw state); rho rows as the reference. */
int or_init_rows(or_ctx *c, const double *F11, const double *F21,
                 const double *F22);

/* Benchmark set-up: full window of free propagators (see kb_oracle.c). */
int or_fill_free_window(or_ctx *c, int tmax, const double *n, double mass);

/* One KB step, evolve! (Evolution.jl:924-981).  t must be tmax+1. */
int or_evolve(or_ctx *c, int t);

/* expandSimData! only (SimDataStruct.jl:262-278). */
int or_expand(or_ctx *c);
int or_window(or_ctx *c, int *tmin, int *tmax);

/* Pieces of evolve!, exposed for kernel-level tests. */
int or_selfenergies(or_ctx *c, int tmone);          /* Evolution.jl:16-31 */
int or_selfenergies_direct(or_ctx *c, int tmone);   /* brute-force double
                                                       convolution twin */
double or_hartree(or_ctx *c, int i, int j);          /* Initialization.jl:207-218 */
/* NinverseNLO pieces: the resummed kernels Irx (calcIrs!, Evolution.jl:
 * 170-206) and IFx (calcIFs!, :354-436) for every tp at tmone; simple = 1
 * runs calcIrssimple!/calcIFssimple! (:294-314, :599-624) instead.
 * or_ninv_selfenergies also evaluates calcSelfEnergyattp! (:683-698). */
int or_ninv_resum(or_ctx *c, int tmone, int simple);
int or_ninv_selfenergies(or_ctx *c, int tmone, int simple);
/* IFx (which = 0) or Irx (which = 1) at the slot of time tp: V complex
 * values, (re, im) interleaved, full grid (the Hermitian extension of the
 * reference's rfft-layout array). */
int or_get_ninv(or_ctx *c, int which, int tp, double *out);
/* Full-lattice transforms the 1/N stage has run so far (for scaling a timed
 * sample to another window length). */
long long or_ninv_transforms(or_ctx *c);
int or_set_hartree(or_ctx *c, double m2);
double or_get_hartree(or_ctx *c);

/* Physical F(i,j) / rho(i,j) lattices (which = 0 F, 1 rho). */
int or_get(or_ctx *c, int which, int i, int j, double *out);
int or_set(or_ctx *c, int which, int i, int j, const double *in);
/* Sigma lattices (which = 0 SigF, 1 Sigr) at slot of time tp. */
int or_get_sigma(or_ctx *c, int which, int tp, double *out);
int or_set_sigma(or_ctx *c, int which, int tp, const double *in);

/* createMeasurement (Measurement.jl:31-99) evaluated per site (every site
 * is its own level with degeneracy 1).  scal[6] = {time, E, E_pot, E_kin,
 * E_mem, M2}.  Per-site arrays (may be NULL): F, ddF, r, n, omega, c.
 * nstab receives the number of sites where the leapfrog stability clamp
 * (Measurement.jl:48-53) fired. */
int or_measure(or_ctx *c, int t, double *scal, double *F, double *ddF,
               double *r, double *n, double *omega, double *cc, int *nstab);

/* Complex 3D DFT on the full grid (unnormalised), sign -1 forward,
 * +1 backward; data interleaved re,im, x fastest.  For tests. */
int or_dft(or_ctx *c, double *data, int sign);

#ifdef __cplusplus
}
#endif
#endif
