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

/* ---- updates ---- */
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
/* FP64 reunitarisation of every link (no reference counterpart: the FP32 form applies it after each step) */
int tcym_reunitarize(tcym_ctx *ctx, int u);

/* ---- observables ---- */
/* ImprovedOperators::ComputeEDotB (improved = 1, CALC/ChernSimonsOperators.cpp:72-388) or
   UnimprovedOperators::ComputeEDotB (improved = 0, :4-53) */
int tcym_edotb(tcym_ctx *ctx, int u, int e, int improved, double *out);
/* ChernSimonsNumber::NCsDot = E.B/(8 pi^2) (TOPOLOGY/ChernSimonsDerivative.cpp:31-49) */
int tcym_ncsdot(tcym_ctx *ctx, int u, int e, double *out);
/* Observables::EdotBMap::CreateMap: per-site improved E.B, map[site] (OBSERVABLES/EdotBMap.cpp:20-348) */
int tcym_edotb_map(tcym_ctx *ctx, int u, int e, double *map);
/* Observables::Bulk::Update: out = {ESqr[3], BSqr[3]} (OBSERVABLES/BulkObservables.cpp:133-143) */
int tcym_bulk(tcym_ctx *ctx, int u, int e, double out[6]);
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
