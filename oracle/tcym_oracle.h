/*
 * tcym_oracle.h -- CPU FP64 restatement of the real-time lattice update loop
 * of TopologicalCYMSolver (leapfrog, Langevin, gradient flow, cooling, the
 * improved E.B Chern-Simons rate, energy, Gauss-law and unitarity violation,
 * Gauss-law restoration and the cooled-trajectory geodesic Delta N_CS).
 *
 * THIS IS TEST INFRASTRUCTURE.  It is the checker that the HIP library is
 * compared against, and the timed CPU baseline.  Nothing in the product path
 * may link, load or call it.
 *
 * Data layout is the reference's own array-of-structures layout:
 *   links  U[s_U * (mu + 3 * site)]            (LATTICE/GaugeLinks.cpp:52-61)
 *   fields E[a + n_E * (mu + 3 * site)]         (LATTICE/ElectricFields.cpp:30-41)
 *   site = x + Nx * (y + Ny * z), periodic wrap  (LATTICE/GaugeLinks.cpp:52-56)
 * SU(2): s_U = 4 reals (u0,u1,u2,u3), U = u3 + i u_a sigma_a
 *        (SUNc/SU2/GroupOperations.cpp:4-9,213-225); n_E = 3.
 * SU(3): s_U = 18 reals = 9 complex (re,im), column-major index i+3j
 *        (SUNc/SU3/GroupOperations.cpp:4-14,28-30); n_E = 8 (Gell-Mann order).
 */
#ifndef TCYM_ORACLE_H
#define TCYM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nc;            /* 2 or 3 */
    int N[3];          /* lattice extents */
    double a[3];       /* lattice spacings (LATTICE/3DGrid.cpp:15) */
    double aScale;     /* LATTICE/3DGrid.cpp:18 */
    int nthreads;      /* OpenMP threads for the sweeps (<=0: runtime default) */
} or_params;

/* sizes */
int  or_su(int nc);            /* reals per link */
int  or_ne(int nc);            /* algebra components */

/* ---- single-matrix operations (exported for known-answer tests) ---- */
void or_mul(int nc, const double *A, const double *B, double *C);          /* C = A B   */
void or_mul_nd(int nc, const double *A, const double *B, double *C);       /* C = A B^+ */
void or_mul_dn(int nc, const double *A, const double *B, double *C);       /* C = A^+ B */
void or_proj(int nc, double c, const double *M, double *out);               /* c ReTr(i t^a M) */
void or_expi(int nc, double c, const double *alpha, double *U);             /* exp(i c alpha^a t^a) */
void or_logi(int nc, double c, const double *U, double *alpha);             /* exp(i c alpha t) = U */
void or_adj_act(int nc, const double *V, const double *E, double *out);    /* (V E V^+)^a */
void or_inv_adj_act(int nc, const double *V, const double *E, double *out);/* (V^+ E V)^a */
double or_unitarity_norm(int nc, const double *U);
void or_reunitarize(int nc, double *U);

/* ---- Philox-4x32-10 and seeded fields ---- */
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void or_normal_pair(const uint32_t r[4], double *z0, double *z1);
/* n normals for (stream, step, link) -- see or_fill_normals in the .c file */
void or_link_normals(uint64_t seed, uint32_t stream, uint32_t step, uint32_t link, int n, double *z);
void or_random_links(const or_params *p, uint64_t seed, double sigma, double *U);
void or_random_efield(const or_params *p, uint64_t seed, double sigma, double *E);
void or_noise_field(const or_params *p, uint64_t seed, uint32_t step, double stddev, double *Xi);
void or_thermal_efield(const or_params *p, uint64_t seed, uint32_t cycle, double beta, double *E);
int  or_thermalize(const or_params *p, double *U, double *E, int cycles, double max_time, double dt, double beta,
                   uint64_t seed, int maxiter, double tol, double *log, int log_rows);

/* ---- lattice sweeps ---- */
void or_leapfrog(const or_params *p, double *U, double *E, int nsteps, double dt);
void or_update_efield(const or_params *p, const double *U, double *E, double dt);   /* E += gamma F */
void or_update_links(const or_params *p, double *U, const double *E, double dt);   /* U <- exp(i cU E) U */
void or_langevin_xi(const or_params *p, double *U, double *E, const double *Xi,
                    double dt, double damping);
void or_langevin(const or_params *p, double *U, double *E, int nsteps, double dt,
                 double damping, double T, uint64_t seed, uint32_t step0);
void or_flow_force(const or_params *p, const double *U, double *E);
void or_flow_links(const or_params *p, double *U, const double *E, double dtau);
void or_flow(const or_params *p, double *U, double *E, int nsteps, double dtau);
double or_flow_calibration(const or_params *p, double *U, double *E, int nsteps, double dtau);

double or_edotb(const or_params *p, const double *U, const double *E);
double or_edotb_unimproved(const or_params *p, const double *U, const double *E);
void   or_edotb_map(const or_params *p, const double *U, const double *E, double *map);
double or_ncsdot(const or_params *p, const double *U, const double *E);
void   or_bulk(const or_params *p, const double *U, const double *E, double out[6]);
/* Block(GaugeLinks**): coarse links U_mu(2x) U_mu(2x+mu) on N/2, 2a (LATTICE/GaugeLinks.cpp:125-151) */
void   or_block_links(const or_params *p, const double *U, double *Uc);
/* Observables::HardScales H[3] (OBSERVABLES/HardScales.cpp:40-104) */
void   or_hard_scales(const or_params *p, const double *U, double out[3]);
void   or_gauss(const or_params *p, const double *U, const double *E, double *smax, double *rms);
void   or_gauss_field(const or_params *p, const double *U, const double *E, double *G);
void   or_unitarity(const or_params *p, const double *U, double *umax, double *rms);
int    or_gauss_restore(const or_params *p, const double *U, double *E, int maxiter,
                        double tol, double *final_max);
void   or_geodesic(const or_params *p, const double *Uold, const double *Unew, double c,
                   double *Umid, double *Emid);
double or_delta_ncs_cooling(const or_params *p, const double *Uold, const double *Unew,
                            double *Umid, double *Emid);
double or_vacuum_ncs(const or_params *p, const double *U);
void   or_reunitarize_field(const or_params *p, double *U);

#ifdef __cplusplus
}
#endif
#endif
