/*
 * tcym.h -- C ABI of the MI355X (gfx950) implementation of the real-time
 * lattice update loop of TopologicalCYMSolver: leapfrog, Langevin and
 * gradient-flow updates, cooling and the cooled-trajectory Delta N_CS, the
 * improved E.B Chern-Simons rate and map, energy, Gauss-law and unitarity
 * violation, and Gauss-law restoration.
 *
 * The reference has no FFI: its "API" is free C++ functions over the
 * GaugeLinks / ElectricFields containers, compiled by textual inclusion.
 * Each entry point below names the reference function it replaces; the
 * drop-in header tcym_dropin.hpp re-exposes those functions with their
 * original names and signatures on top of this ABI.
 *
 * Conventions
 *  - every call returns an int status (TCYM_OK = 0, negative on error) and
 *    never exits; tcym_last_error() gives a message;
 *  - host arrays are owned by the caller and never retained;
 *  - host field arrays use the reference's own array-of-structures layout:
 *      links  U[s_U*(mu + 3*site)]   s_U = 4 (SU(2) quaternion u0,u1,u2,u3,
 *             U = u3 + i u_a sigma_a) or 18 (SU(3), 9 complex column-major
 *             element (i,j) at i+3j, stored re,im)   (LATTICE/GaugeLinks.cpp:52-61)
 *      fields E[a + n_E*(mu + 3*site)]   n_E = 3 or 8   (LATTICE/ElectricFields.cpp:30-41)
 *      site = x + Nx*(y + Ny*z)
 *  - fields live on the device in numbered slots (see TCYM_U_* / TCYM_E_*);
 *    a slot is allocated on first use and keeps its contents between calls;
 *  - calls are synchronous at the ABI (results are valid on return);
 *  - one context per host thread; a context is not thread-safe.
 */
#ifndef TCYM_H
#define TCYM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TCYM_OK               0
#define TCYM_E_ARG           -1   /* bad argument (slot, size, group, ...) */
#define TCYM_E_HIP           -2   /* HIP runtime error */
#define TCYM_E_NOCONV        -3   /* Gauss-law restoration hit its iteration cap */
#define TCYM_E_NONFINITE     -4   /* a reduction produced NaN/Inf */
#define TCYM_E_UNSUPPORTED   -5   /* device is not gfx950, or feature not built */
#define TCYM_E_IO            -6   /* a file could not be opened, read or written, or is malformed (tcym_io.h) */

#define TCYM_FP64 0               /* FP64 storage and arithmetic */
#define TCYM_FP32 1               /* FP32 storage/arithmetic; FP64 reunitarisation,
                                     FP64 sums and FP64 Gauss-law restoration */

/* link slots */
#define TCYM_U_HOT  0             /* GLinks::U */
#define TCYM_U_COOL 1             /* Cooling::DynamicFields::U */
#define TCYM_U_OLD  2             /* ChernSimonsNumber::CoolingMethod::UOld */
#define TCYM_U_NEW  3             /* ChernSimonsNumber::CoolingMethod::UNew */
#define TCYM_U_MID  4             /* ChernSimonsNumber::CoolingMethod::UMid */
#define TCYM_U_AUX  5             /* free for callers */
#define TCYM_NUM_U  6
/* E-field slots */
#define TCYM_E_HOT  0             /* EFields::E */
#define TCYM_E_COOL 1             /* Cooling::DynamicFields::E */
#define TCYM_E_MID  2             /* ChernSimonsNumber::CoolingMethod::EMid */
#define TCYM_E_XI   3             /* LangevinVariables::Xi */
#define TCYM_E_AUX  4             /* free for callers */
#define TCYM_NUM_E  5

typedef struct tcym_ctx tcym_ctx;

typedef struct {
    int nc;            /* 2 or 3: SU_Nc_FLAG (SUNc/Definitions.cpp:103-131) as a run-time value */
    int precision;     /* TCYM_FP64 or TCYM_FP32 */
    int N[3];          /* Lattice::N (LATTICE/3DGrid.cpp:14) */
    double a[3];       /* Lattice::a (LATTICE/3DGrid.cpp:15) */
    double aScale;     /* Lattice::aScale (LATTICE/3DGrid.cpp:18) */
    int device;        /* HIP device ordinal */
} tcym_params;

/* ---- context ---- */
int tcym_create(const tcym_params *p, tcym_ctx **out);      /* Lattice::Init (LATTICE/GLinksEFields.cpp:36-45) */
int tcym_destroy(tcym_ctx *ctx);
const char *tcym_last_error(const tcym_ctx *ctx);
/* the parameters the context was created with (no reference counterpart:
   the reference reads the Lattice namespace globals) */
int tcym_get_params(const tcym_ctx *ctx, tcym_params *p);
const char *tcym_version(void);

/* ---- fields ---- */
int tcym_set_links(tcym_ctx *ctx, int slot, const double *U_aos);   /* fill GaugeLinks::U from host */
int tcym_get_links(tcym_ctx *ctx, int slot, double *U_aos);
int tcym_set_efield(tcym_ctx *ctx, int slot, const double *E_aos);  /* fill ElectricFields::E from host */
int tcym_get_efield(tcym_ctx *ctx, int slot, double *E_aos);
int tcym_copy_links(tcym_ctx *ctx, int dst, int src);               /* Copy(GaugeLinks*,GaugeLinks*) (LATTICE/GaugeLinks.cpp:116-132) */
int tcym_copy_efield(tcym_ctx *ctx, int dst, int src);              /* Copy(ElectricFields*,ElectricFields*) (LATTICE/ElectricFields.cpp:90-107) */
int tcym_zero_efield(tcym_ctx *ctx, int slot);                      /* ElectricFields::SetZero (LATTICE/ElectricFields.cpp:48-63) */
int tcym_unit_links(tcym_ctx *ctx, int slot);                       /* GaugeLinks::SetIdentity (LATTICE/GaugeLinks.cpp:74-87) */
/* seeded initial conditions (Philox-4x32-10; counter (link, step, stream, j)):
   U = exp(i eps^a t^a), eps^a ~ N(0, sigma^2) per (site, mu, a)   [stream 0]
   E^a ~ N(0, sigma^2)                                            [stream 1]
   Replaces INITIALCONDITIONS/SetRandomMatrices.cpp:3-56 and the GSL draws of
   MISC/RNG/GSLRandomNumberGenerator.cpp:31-62 for the seeded configurations. */
int tcym_random_links(tcym_ctx *ctx, int slot, uint64_t seed, double sigma);
int tcym_random_efield(tcym_ctx *ctx, int slot, uint64_t seed, double sigma);

/* ---- updates ----
   An update call of nsteps steps is faster than nsteps calls of one step:
   between its steps the SU(3) kernels keep only columns 0 and 1 of the
   links (the next step rebuilds column 2), and the host synchronises once.
   The fields are bit-identical either way, and the links a call leaves in
   its slot are complete. */
/* Dynamics::Update x nsteps with dTau = dt (DYNAMICS/Dynamics.cpp:204-218) */
int tcym_leapfrog(tcym_ctx *ctx, int u, int e, int nsteps, double dt);
/* Dynamics::UpdateElectricFields / UpdateGaugeLinks on the full box (DYNAMICS/Dynamics.cpp:145,97) */
int tcym_update_efield(tcym_ctx *ctx, int u, int e, double dt);
int tcym_update_links(tcym_ctx *ctx, int u, int e, double dt);
/* LangevinDynamics::Update x nsteps: Xi ~ N(0, 2 damping T) drawn in-kernel
   from Philox (stream 2, counter step0 + k), then UpdateFields with damping
   (DYNAMICS/LangevinDynamics.cpp:260-347; the reference has damping 1, T = 1/beta) */
int tcym_langevin(tcym_ctx *ctx, int u, int e, int nsteps, double dt, double damping, double T,
                  uint64_t seed, uint32_t step0);
/* Hot-trajectory N_CS: nsteps of leapfrog (mode 0) or Langevin (mode 1, as
   tcym_langevin with seed and step0), adding to *dncs after every step
   dt/2 [NCsDot(U_n, E_n+1/2) + NCsDot(U_n+1, E_n+1/2)].  The reference
   declares this integral (ChernSimonsNumber::DeltaNCsRealTime,
   TOPOLOGY/ChernSimonsDerivative.cpp:8) and resets it
   (SphaleronSimulation.cpp:405) but never accumulates it; the rule mirrors
   the calibration flow's (GRADIENTFLOW/GradientFlow.cpp:168-188).  The sum
   stays on the device until the end of the call. */
int tcym_evolve_ncs(tcym_ctx *ctx, int mode, int u, int e, int nsteps, double dt, double damping, double T,
                    uint64_t seed, uint32_t step0, double *dncs);
/* LangevinDynamics::UpdateFields with the noise slot xi (DYNAMICS/LangevinDynamics.cpp:300-311) */
int tcym_langevin_xi(tcym_ctx *ctx, int u, int e, int xi, double dt, double damping);
/* LangevinDynamics::UpdateRandomGaussianNoise into slot xi (DYNAMICS/LangevinDynamics.cpp:260-286) */
int tcym_noise(tcym_ctx *ctx, int xi, uint64_t seed, uint32_t step, double stddev);
/* GradientFlow::StandardUpdate (calibrate = 0) or CalibrationUpdate (calibrate = 1,
   *dncs += accumulated Delta N_CS) x nsteps, SqrtDcTauOverSpacing^2 = dtau
   (GRADIENTFLOW/GradientFlow.cpp:150-188); e < 0 runs the fused flow without
   storing the force field */
int tcym_flow(tcym_ctx *ctx, int u, int e, int nsteps, double dtau, int calibrate, double *dncs);
/* Cooling::CoolNSave: copy u_hot into the cooling slots, zero E, flow nsteps,
   store the cooled links in u_cold (GRADIENTFLOW/Cooling.cpp:156-179) */
int tcym_cool(tcym_ctx *ctx, int u_hot, int u_cold, int nsteps, double dtau);
/* ChernSimonsNumber::CoolingMethod::Start / Update(Measure) (TOPOLOGY/CoolingMethod.cpp:127-165):
   cool u_hot into TCYM_U_NEW; if measure, *dncs = GetDeltaNCs(1) between
   TCYM_U_OLD and TCYM_U_NEW; then TCYM_U_OLD <- TCYM_U_NEW */
int tcym_cooling_start(tcym_ctx *ctx, int u_hot, int nsteps, double dtau);
int tcym_cooling_update(tcym_ctx *ctx, int u_hot, int nsteps, double dtau, int measure, double *dncs);

/* ---- blocking and calibration ---- */
/* Block(GaugeLinks**) / Block(ElectricFields**) (LATTICE/GaugeLinks.cpp:125-151,
   LATTICE/ElectricFields.cpp:113-127): the links of slot u become the lattice
   of half the extent and twice the spacing, U'_mu(x) = U_mu(2x) U_mu(2x+mu);
   E slot e (if >= 0) becomes zero at the new resolution.  A slot keeps its
   resolution level (0 = the full lattice) until it is overwritten; every entry
   point works at the level of the slots it is given (mixing levels is an
   error), and a level-l slot reads and writes host arrays of the level-l size. */
int tcym_block(tcym_ctx *ctx, int u, int e);
/* resolution of link slot u or (u < 0) field slot e: blocking level and extents */
int tcym_level(tcym_ctx *ctx, int u, int e, int *level, int N[3]);
/* Cooling::CoolNSave with blocking (GRADIENTFLOW/Cooling.cpp:69-179): u_cold := u_hot,
   then nsteps fused flow steps; after step k (from 1) the cold links are blocked
   when k % (block_frequency 8^b) == 0 and b < max_blocking, b the blockings done
   so far (block_frequency 0: never).  *nblocked (may be NULL) receives b. */
int tcym_cool_blocked(tcym_ctx *ctx, int u_hot, int u_cold, int nsteps, double dtau, int block_frequency,
                      int max_blocking, int *nblocked);
/* Cooling::Perform with Calibrate = 0 (GRADIENTFLOW/Cooling.cpp:69-147) in place on link
   slot u: flow steps step0+1 .. nsteps, blocked as in tcym_cool_blocked with the blocking
   count continuing from *nblocked (in/out).  tcym_cool_blocked is a copy followed by this
   call from step 0; a caller that checks the cooled links every so many steps (as
   Cooling::Perform does every 100, :113-119) continues the cooling in pieces. */
int tcym_cool_continue(tcym_ctx *ctx, int u, int step0, int nsteps, double dtau, int block_frequency,
                       int max_blocking, int *nblocked);
/* Cooling::ContinueCooling with GradientFlow::CalibrationUpdate (GRADIENTFLOW/Cooling.cpp:181-195,
   69-147; GradientFlow.cpp:168-188): u_work := u_src with E slot e_work zero, then the
   calibration flow steps step0+1 .. nsteps (the reference's CoolingSteps continue from
   the preceding cooling), blocking as in tcym_cool_blocked with the blocking count
   continuing from *nblocked (in/out).  *dncs (may be NULL) receives DeltaNCsCooling
   along these steps; log (may be NULL) receives nsteps - step0 + 1 rows of
   {CoolingTime, Bulk::MAGNETIC, Bulk::ELECTRIC, DeltaNCsCooling, VacuumEstimator::NCS}
   before the first step and after each step (Cooling.cpp:85-111), the cooling time
   starting at ctau0 (GradientFlow::cTau). */
int tcym_continue_cooling(tcym_ctx *ctx, int u_src, int u_work, int e_work, int step0, int nsteps, double dtau,
                          double ctau0, int block_frequency, int max_blocking, int *nblocked, double *dncs,
                          double *log);
/* ---- thermal initial condition ---- */
/* ThermalDynamics::Thermalize(MaxSteps = cycles, MaxTime) (DYNAMICS/ThermalDynamics.cpp:260-316):
   per cycle i, E^a ~ N(0, 1/beta) (Philox stream 3, counter step i, replacing
   RandomNumberGenerator::Gauss()/sqrt(beta), :205-226) and Gauss-law
   restoration (maxiter, tol as tcym_gauss_restore), then leapfrog steps of dt
   (the reference's dTau = 0.0375) while steps*dt < max_time.  Every 20 steps a
   row {Time() + i max_time, T00, TXX, TYY, TZZ, ELECTRIC, MAGNETIC} of the
   reference's ThermalEnergy file goes to log (log_rows rows of 7 doubles at
   most; *nrows receives the number of rows produced).  The links are evolved
   from their current values (the reference starts from InitialConditions::SetZero). */
int tcym_thermalize(tcym_ctx *ctx, int u, int e, int cycles, double max_time, double dt, double beta, uint64_t seed,
                    int maxiter, double tol, double *log, int log_rows, int *nrows);
/* InitialConditions::SetQuasiParticles without its final checks and Gauss-law
   restoration (INITIALCONDITIONS/SetQuasiParticles.cpp:24-174, Minkowski metric):
   momentum modes A_mu(p), E_mu(p) of occupancy f(p) = n0/sqrt(p^2/Qs^2 + 0.1) for
   |p| < Qs along the free-field polarisations (MISC/MinkowskiWaveVectors.cpp),
   an inverse FFT (rocFFT in place of FFTW), then U_mu = exp(-i A_mu^a t^a) and E_mu
   from the real parts into link slot u and field slot e.  The complex normals of
   the modes come from Philox stream 4 (counter colour + NE p, key seed) in place
   of RandomNumberGenerator::ComplexGauss.  Follow with tcym_gauss_restore, as
   the reference does. */
int tcym_quasi_particles(tcym_ctx *ctx, int u, int e, double Qs, double n0, uint64_t seed);
/* FP64 reunitarisation of every link (no reference counterpart: the FP32 form applies it after each step) */
int tcym_reunitarize(tcym_ctx *ctx, int u);

/* ---- observables ---- */
/* ImprovedOperators::ComputeEDotB (improved = 1, CALC/ChernSimonsOperators.cpp:72-388) or
   UnimprovedOperators::ComputeEDotB (improved = 0, :4-53) */
int tcym_edotb(tcym_ctx *ctx, int u, int e, int improved, double *out);
/* ChernSimonsNumber::NCsDot = E.B/(8 pi^2) (TOPOLOGY/ChernSimonsDerivative.cpp:31-49) */
int tcym_ncsdot(tcym_ctx *ctx, int u, int e, double *out);
/* tcym_ncsdot without any allocation.  No counterpart in the reference: it is
   for callers that cannot allocate device memory (code under HIP graph
   capture, serving loops, harnesses that forbid hipMalloc).  out is DEVICE
   memory of one double, written asynchronously on the context's stream
   (tcym_sync waits); the value is bit-identical to tcym_ncsdot's.  Slots u and
   e must already hold fields (this call never allocates a slot).  The block
   partials live in a fixed 1 MiB device array per GPU (131072 partial sums:
   lattices up to about 220^3); larger lattices are refused with TCYM_E_ARG.
   Every call on a GPU shares that array: calls that overlap on one GPU (two
   contexts on two host threads) race. */
int tcym_ncsdot_noalloc(tcym_ctx *ctx, int u, int e, double *out);
/* Observables::EdotBMap::CreateMap: per-site improved E.B, map[site] (OBSERVABLES/EdotBMap.cpp:20-348) */
int tcym_edotb_map(tcym_ctx *ctx, int u, int e, double *map);
/* Observables::Bulk::Update: out = {ESqr[3], BSqr[3]} (OBSERVABLES/BulkObservables.cpp:133-143) */
int tcym_bulk(tcym_ctx *ctx, int u, int e, double out[6]);
/* Observables::HardScales::Update: H[mu] = sum_x sum_a (2 a^3 F_mu^a)^2 / V,
   F the force of the E-field update (OBSERVABLES/HardScales.cpp:40-104);
   LambdaXX/YY/ZZ are combinations of H on the host (:26-36) */
int tcym_hard_scales(tcym_ctx *ctx, int u, double out[3]);
/* Observables::GaussLaw::UpdateViolation: signed max and RMS (OBSERVABLES/GaussLawViolation.cpp:13-61) */
int tcym_gauss(tcym_ctx *ctx, int u, int e, double *signed_max, double *rms);
/* Observables::Unitarity::UpdateViolation (OBSERVABLES/UnitarityViolation.cpp:13-51) */
int tcym_unitarity(tcym_ctx *ctx, int u, double *max, double *rms);
/* GaussLawRestoration::Restore with iteration cap and tolerance (GAUSSLAWFIXING/GaussRestore.cpp:131-167;
   the reference uses maxiter 10000, tol 1e-14); returns TCYM_E_NOCONV when the cap is hit */
int tcym_gauss_restore(tcym_ctx *ctx, int u, int e, int maxiter, double tol, int *iters, double *final_max);
/* CoolingMethod::PerformInterpolation(c) (TOPOLOGY/CoolingMethod.cpp:37-68) */
int tcym_geodesic(tcym_ctx *ctx, int uold, int unew, int umid, int emid, double c);
/* CoolingMethod::GetDeltaNCs(1) (TOPOLOGY/CoolingMethod.cpp:80-121) */
int tcym_delta_ncs_cooling(tcym_ctx *ctx, int uold, int unew, double *out);
/* ChernSimonsNumber::VacuumEstimator::NCS (TOPOLOGY/VacuumEstimator.cpp:7-86) */
int tcym_vacuum_ncs(tcym_ctx *ctx, int u, double *out);

/* ---- update kernel form ---- */
/* Form of the leapfrog, Langevin and flow update kernels (no reference
   counterpart: the reference has one CPU loop).  The plaquette form
   multiplies out each plaquette once and shares it among its four links; it
   takes FP32 storage and lattices with N[0] a multiple of 64 and at most
   384 and N[1] with a divisor PY >= 2, N[0] PY <= 768.  The per-link form
   computes the staples of every link apart.  TCYM_FORM_AUTO (the default)
   runs the plaquette form where it applies and was measured faster (its
   grid fills the device with at least two blocks per compute unit, and for
   SU(3) N[0] >= 192: SU(3) 192^3 and 256^3, SU(2) 128^3 and up), else the
   per-link form; TCYM_FORM_LINK always the per-link form;
   TCYM_FORM_PLAQUETTE the plaquette form wherever it applies.  The two forms
   sum the same force terms in another order (results agree to FP32
   rounding); the batch kernels use the per-link form. */
#define TCYM_FORM_AUTO      0
#define TCYM_FORM_LINK      1
#define TCYM_FORM_PLAQUETTE 2
int tcym_set_update_form(tcym_ctx *ctx, int form);

/* ---- measurement support ---- */
/* per-kernel HIP-event timing on the context's stream */
int tcym_timing(tcym_ctx *ctx, int enable);
int tcym_kernel_time(tcym_ctx *ctx, const char *kernel, double *total_ms, long long *launches);
int tcym_timing_reset(tcym_ctx *ctx);
int tcym_sync(tcym_ctx *ctx);
/* device bytes held by the context */
int tcym_device_bytes(tcym_ctx *ctx, long long *bytes);

#ifdef __cplusplus
}
#endif
#endif
