/* Batches of independent lattices of uneven sizes on one MI355X.
 *
 * The reference runs independent trajectories as separate MPI ranks, each
 * with seed + rank (SU2Solver.cpp:161-164) and its own copy of every
 * sweep.  A batch holds many such lattices in one device allocation and runs
 * each pass (an update step, a flow step, a measurement) over all of them
 * in one persistent kernel, driven by per-lattice offset tables.  Every
 * lattice of a batch computes exactly what a tcym_ctx of the same size,
 * precision and seed computes: fields and per-lattice sums are bit-identical
 * to the single-lattice entry points named below.
 *
 * Slots are the single-lattice slot numbers (TCYM_U_HOT, TCYM_E_HOT, ...);
 * the lattice index e runs over 0 .. count-1.  All entry points return a
 * TCYM_* status and are synchronous.  These entry points are defined by
 * this library; the reference has no batched interface.
 */
#ifndef TCYM_BATCH_H
#define TCYM_BATCH_H

#include <stdint.h>

#include "tcym.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tcym_batch tcym_batch;

typedef struct {
    int nc;            /* 2 or 3 */
    int precision;     /* TCYM_FP64 or TCYM_FP32 */
    int count;         /* number of lattices */
    const int *N;      /* count x 3 extents, N[3e + k] */
    double a[3];       /* lattice spacings, common to the batch */
    double aScale;
    int device;
} tcym_batch_params;

int tcym_batch_create(const tcym_batch_params *p, tcym_batch **out);
int tcym_batch_destroy(tcym_batch *b);
const char *tcym_batch_last_error(const tcym_batch *b);
int tcym_batch_sync(tcym_batch *b);
int tcym_batch_device_bytes(tcym_batch *b, long long *bytes);

/* Fields of lattice e in the reference's AoS FP64 layout (as tcym_set_links
 * / tcym_get_links / tcym_set_efield / tcym_get_efield). */
int tcym_batch_set_links(tcym_batch *b, int slot, int e, const double *U);
int tcym_batch_get_links(tcym_batch *b, int slot, int e, double *U);
int tcym_batch_set_efield(tcym_batch *b, int slot, int e, const double *E);
int tcym_batch_get_efield(tcym_batch *b, int slot, int e, double *E);

/* Seeded initial condition of every lattice: lattice e gets exactly what an
 * FP64 tcym_ctx gets from tcym_random_links(seed0 + e, sigma_u),
 * tcym_random_efield(seed0 + e, sigma_e) and tcym_gauss_restore(maxiter,
 * tol) (GAUSSLAWFIXING/GaussRestore.cpp:131-167), stored in the slots
 * TCYM_U_HOT / TCYM_E_HOT.  iters (count entries, may be NULL) receives the
 * restoration iterations per lattice. */
int tcym_batch_initial_condition(tcym_batch *b, uint64_t seed0, double sigma_u, double sigma_e,
                                 int maxiter, double tol, int *iters);

/* Philox step counter of the next Langevin step (0 after creation). */
int tcym_batch_set_step(tcym_batch *b, uint32_t step);
int tcym_batch_get_step(tcym_batch *b, uint32_t *step);

/* nsteps leapfrog steps of every lattice (as tcym_leapfrog). */
int tcym_batch_leapfrog(tcym_batch *b, int nsteps, double dt);
/* nsteps Langevin steps of every lattice; lattice e uses key seed0 + e and
 * Philox steps step .. step + nsteps - 1 (as tcym_langevin(seed0 + e, step)). */
int tcym_batch_langevin(tcym_batch *b, int nsteps, double dt, double damping, double T, uint64_t seed0);
/* nsteps fused gradient-flow steps of link slot u of every lattice, the force
 * not stored (as tcym_flow(u, e = -1) / the flow of tcym_cool). */
int tcym_batch_flow(tcym_batch *b, int u, int nsteps, double dtau);

/* Fused update+measure segments captured once in a HIP graph and replayed:
 * each segment runs `steps` updates (mode 0 leapfrog, 1 Langevin with key
 * seed0 + e) of every lattice, then NCsDot (TOPOLOGY/ChernSimonsDerivative.cpp:31-49)
 * of every lattice into ncsdot[e * nseg + j] for segment j (may be NULL).
 * Bit-identical to the same steps and measurements run one call at a time. */
int tcym_batch_evolve(tcym_batch *b, int mode, int nseg, int steps, double dt, double damping, double T,
                      uint64_t seed0, double *ncsdot);

/* CoolingMethod (TOPOLOGY/CoolingMethod.cpp:19-165) on every lattice, as
 * tcym_cooling_start / tcym_cooling_update with u_hot = TCYM_U_HOT;
 * dncs (count entries) receives Delta N_CS of each lattice. */
int tcym_batch_cooling_start(tcym_batch *b, int nsteps, double dtau);
int tcym_batch_cooling_update(tcym_batch *b, int nsteps, double dtau, double *dncs);

/* Per-lattice measurements (count entries per quantity, lattice-major):
 * NCsDot (as tcym_ncsdot), the 6 bulk sums (as tcym_bulk), Gauss-law
 * violation {signed max, rms} (as tcym_gauss) and unitarity violation
 * {max, rms} (as tcym_unitarity). */
int tcym_batch_ncsdot(tcym_batch *b, int u, int e, double *out);
int tcym_batch_bulk(tcym_batch *b, int u, int e, double *out);
int tcym_batch_gauss(tcym_batch *b, int u, int e, double *out);
int tcym_batch_unitarity(tcym_batch *b, int u, double *out);

/* Autotuner of the update kernels: times every candidate shape (per-link
 * form: canonical, linear and 64/32/16-wide patch walks, shapes -1, 0, 4..6;
 * FP32 batches whose rows fit a 256-thread block also the plaquette form with
 * units of 2, 4 or 8 planes, shapes 7..9) and persistent grid size on
 * scratch copies of the fields and keeps the fastest for one pass: mode 0 or
 * 1 tunes the leapfrog and Langevin updates (repeated launches on one
 * input), mode 3 the flow (sequences of flow steps alternating between two
 * link buffers with two-column stores, as in a cooling episode); each pass
 * keeps its own choice.  The timing runs overwrite the free slots
 * TCYM_U_AUX and TCYM_E_AUX.  Within a form results do not depend on the
 * choice (each lattice's fields equal those of a tcym_ctx run in that form,
 * TCYM_FORM_LINK or TCYM_FORM_PLAQUETTE); the two forms agree to FP32
 * rounding.  choice receives {shape, blocks per CU, XCD split} and us the
 * best time per update step in microseconds (both may be NULL). */
int tcym_batch_tune(tcym_batch *b, int mode, double dt, int reps, int choice[3], double *us);

/* The schedule of one pass set by hand (mode 0 or 1: leapfrog and Langevin;
 * mode 3: flow): shape as in tcym_batch_tune, blocks per compute unit of the
 * persistent grid (1..64), XCD ranges on or off.  No reference counterpart
 * (the reference runs one lattice per process); for tests and for callers
 * that keep a tuned schedule across runs.  TCYM_E_UNSUPPORTED for a
 * plaquette shape on an FP64 batch or on rows that do not fit its blocks. */
int tcym_batch_set_schedule(tcym_batch *b, int mode, int shape, int blocks_per_cu, int xcd_split);

/* Kernel timing as tcym_timing / tcym_kernel_time / tcym_timing_reset. */
int tcym_batch_timing(tcym_batch *b, int enable);
int tcym_batch_kernel_time(tcym_batch *b, const char *name, double *total_ms, long long *launches);
int tcym_batch_timing_reset(tcym_batch *b);

#ifdef __cplusplus
}
#endif

#endif
