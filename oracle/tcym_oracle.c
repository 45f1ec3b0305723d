/*
 * tcym_oracle.c -- CPU FP64 restatement (the oracle).  TEST INFRASTRUCTURE:
 * only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may load
 * it, and only as the checker.  See tcym_oracle.h for the data layout.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "tcym_oracle.h"
#include "or_su2.h"
#include "or_su3.h"

/* ---- group-specialised lattice code ---- */
#define GP(n) su2_##n
#define LF(n) su2l_##n
#define GSU SU2_SU
#define GNE SU2_NE
#include "or_lattice.inc"
#undef GP
#undef LF
#undef GSU
#undef GNE
#undef FOR_SITES

#define GP(n) su3_##n
#define LF(n) su3l_##n
#define GSU SU3_SU
#define GNE SU3_NE
#include "or_lattice.inc"
#undef GP
#undef LF
#undef GSU
#undef GNE
#undef FOR_SITES

int or_su(int nc) { return nc == 2 ? SU2_SU : SU3_SU; }
int or_ne(int nc) { return nc == 2 ? SU2_NE : SU3_NE; }

static long vol(const or_params *p) { return (long)p->N[0]*p->N[1]*p->N[2]; }

static void set_threads(const or_params *p) {
#ifdef _OPENMP
    if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#else
    (void)p;
#endif
}

#define DISPATCH(p, call2, call3) do { set_threads(p); if ((p)->nc == 2) { call2; } else { call3; } } while (0)

/* ---- matrix-level entry points ---- */
void or_mul(int nc, const double *A, const double *B, double *C) { if (nc == 2) su2_mul(A, B, C); else su3_mul(A, B, C); }
void or_mul_nd(int nc, const double *A, const double *B, double *C) { if (nc == 2) su2_mul_nd(A, B, C); else su3_mul_nd(A, B, C); }
void or_mul_dn(int nc, const double *A, const double *B, double *C) { if (nc == 2) su2_mul_dn(A, B, C); else su3_mul_dn(A, B, C); }
void or_proj(int nc, double c, const double *M, double *o) { if (nc == 2) su2_proj(c, M, o); else su3_proj(c, M, o); }
void or_expi(int nc, double c, const double *al, double *U) { if (nc == 2) su2_expi(c, al, U); else su3_expi(c, al, U); }
void or_logi(int nc, double c, const double *U, double *al) { if (nc == 2) su2_logi(c, U, al); else su3_logi(c, U, al); }
void or_adj_act(int nc, const double *V, const double *E, double *o) { if (nc == 2) su2_adj_act(V, E, o); else su3_adj_act(V, E, o); }
void or_inv_adj_act(int nc, const double *V, const double *E, double *o) { if (nc == 2) su2_inv_adj_act(V, E, o); else su3_inv_adj_act(V, E, o); }
double or_unitarity_norm(int nc, const double *U) { return nc == 2 ? su2_unitarity_norm(U) : su3_unitarity_norm(U); }
void or_reunitarize(int nc, double *U) { if (nc == 2) su2_reunitarize(U); else su3_reunitarize(U); }

/* ---- Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ---- */
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u*c0, p1 = (uint64_t)0xCD9E8D57u*c2;
        uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0;
        uint32_t h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
        c0 = h1 ^ c1 ^ k0; c1 = l1; c2 = h0 ^ c3 ^ k1; c3 = l0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* two uint32 -> uniform in (0,1] with 53 random bits; Box-Muller pair */
static double u53(uint32_t hi, uint32_t lo) {
    uint64_t x = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
    return ((double)x + 1.0)*(1.0/9007199254740992.0);
}
void or_normal_pair(const uint32_t r[4], double *z0, double *z1) {
    double u1 = u53(r[0], r[1]), u2 = u53(r[2], r[3]);
    double R = sqrt(-2.0*log(u1)), t = 6.283185307179586476925286766559*u2;
    *z0 = R*cos(t); *z1 = R*sin(t);
}
/* normals for one link: counter (link, step, stream, j), key = seed halves */
void or_link_normals(uint64_t seed, uint32_t stream, uint32_t step, uint32_t link, int n, double *z) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int j = 0; 2*j < n; j++) {
        uint32_t ctr[4] = {link, step, stream, (uint32_t)j}, r[4];
        double a, b;
        or_philox4x32(ctr, key, r);
        or_normal_pair(r, &a, &b);
        z[2*j] = a;
        if (2*j + 1 < n) z[2*j + 1] = b;
    }
}

/* U = exp(i eps^a t^a), eps^a ~ N(0, sigma^2), stream 0 */
void or_random_links(const or_params *p, uint64_t seed, double sigma, double *U) {
    int su = or_su(p->nc), ne = or_ne(p->nc);
    long L = 3*vol(p);
    set_threads(p);
    #pragma omp parallel for schedule(static)
    for (long l = 0; l < L; l++) {
        double z[8];
        or_link_normals(seed, 0u, 0u, (uint32_t)l, ne, z);
        for (int a = 0; a < ne; a++) z[a] *= sigma;
        or_expi(p->nc, 1.0, z, U + (size_t)su*l);
    }
}
/* E^a ~ N(0, sigma^2), stream 1 */
void or_random_efield(const or_params *p, uint64_t seed, double sigma, double *E) {
    int ne = or_ne(p->nc);
    long L = 3*vol(p);
    set_threads(p);
    #pragma omp parallel for schedule(static)
    for (long l = 0; l < L; l++) {
        double z[8];
        or_link_normals(seed, 1u, 0u, (uint32_t)l, ne, z);
        for (int a = 0; a < ne; a++) E[(size_t)ne*l + a] = sigma*z[a];
    }
}
/* Langevin noise Xi^a ~ N(0, stddev^2), stream 2, counter step */
void or_noise_field(const or_params *p, uint64_t seed, uint32_t step, double stddev, double *Xi) {
    int ne = or_ne(p->nc);
    long L = 3*vol(p);
    set_threads(p);
    #pragma omp parallel for schedule(static)
    for (long l = 0; l < L; l++) {
        double z[8];
        or_link_normals(seed, 2u, step, (uint32_t)l, ne, z);
        for (int a = 0; a < ne; a++) Xi[(size_t)ne*l + a] = stddev*z[a];
    }
}

/* ThermalDynamics::ReassignElectricFields: E^a ~ N(0, 1/beta), stream 3 with
   the cycle as counter step (DYNAMICS/ThermalDynamics.cpp:205-221; Philox in
   place of RandomNumberGenerator::Gauss) */
void or_thermal_efield(const or_params *p, uint64_t seed, uint32_t cycle, double beta, double *E) {
    int ne = or_ne(p->nc);
    long L = 3*vol(p);
    const double sigma = 1.0/sqrt(beta);
    set_threads(p);
    #pragma omp parallel for schedule(static)
    for (long l = 0; l < L; l++) {
        double z[8];
        or_link_normals(seed, 3u, cycle, (uint32_t)l, ne, z);
        for (int a = 0; a < ne; a++) E[(size_t)ne*l + a] = sigma*z[a];
    }
}

/* ---- lattice entry points ---- */
void or_leapfrog(const or_params *p, double *U, double *E, int n, double dt) {
    DISPATCH(p, su2l_leapfrog(p, U, E, n, dt), su3l_leapfrog(p, U, E, n, dt));
}
void or_update_efield(const or_params *p, const double *U, double *E, double dt) {
    DISPATCH(p, su2l_update_efield(p, U, E, dt), su3l_update_efield(p, U, E, dt));
}
void or_update_links(const or_params *p, double *U, const double *E, double dt) {
    DISPATCH(p, su2l_update_links(p, U, E, dt), su3l_update_links(p, U, E, dt));
}
void or_langevin_xi(const or_params *p, double *U, double *E, const double *Xi, double dt, double damping) {
    DISPATCH(p, su2l_langevin_xi(p, U, E, Xi, dt, damping), su3l_langevin_xi(p, U, E, Xi, dt, damping));
}
/* LangevinDynamics::Update: noise of variance 2 damping T (2/beta at the
   reference's damping 1, DYNAMICS/LangevinDynamics.cpp:260-286), then UpdateFields */
void or_langevin(const or_params *p, double *U, double *E, int n, double dt, double damping,
                 double T, uint64_t seed, uint32_t step0) {
    long L = 3*vol(p);
    double *Xi = (double*)malloc(sizeof(double)*(size_t)L*or_ne(p->nc));
    double sd = sqrt(2.0*damping*T);
    for (int k = 0; k < n; k++) {
        or_noise_field(p, seed, step0 + (uint32_t)k, sd, Xi);
        or_langevin_xi(p, U, E, Xi, dt, damping);
    }
    free(Xi);
}
void or_flow_force(const or_params *p, const double *U, double *E) {
    DISPATCH(p, su2l_flow_force(p, U, E), su3l_flow_force(p, U, E));
}
void or_flow_links(const or_params *p, double *U, const double *E, double dtau) {
    DISPATCH(p, su2l_flow_links(p, U, E, dtau), su3l_flow_links(p, U, E, dtau));
}
/* GradientFlow::StandardUpdate x nsteps (GRADIENTFLOW/GradientFlow.cpp:150-165) */
void or_flow(const or_params *p, double *U, double *E, int n, double dtau) {
    for (int k = 0; k < n; k++) { or_flow_force(p, U, E); or_flow_links(p, U, E, dtau); }
}
double or_edotb(const or_params *p, const double *U, const double *E) {
    double r = 0.0; DISPATCH(p, r = su2l_edotb(p, U, E), r = su3l_edotb(p, U, E)); return r;
}
double or_edotb_unimproved(const or_params *p, const double *U, const double *E) {
    double r = 0.0; DISPATCH(p, r = su2l_edotb_unimproved(p, U, E), r = su3l_edotb_unimproved(p, U, E)); return r;
}
void or_edotb_map(const or_params *p, const double *U, const double *E, double *map) {
    DISPATCH(p, su2l_edotb_map(p, U, E, map), su3l_edotb_map(p, U, E, map));
}
/* ChernSimonsNumber::NCsDot = 0.125 E.B / pi^2 (TOPOLOGY/ChernSimonsDerivative.cpp:31-49) */
double or_ncsdot(const or_params *p, const double *U, const double *E) {
    return 0.125*or_edotb(p, U, E)/(M_PI*M_PI);
}
/* GradientFlow::CalibrationUpdate x nsteps: returns the accumulated
   Delta N_CS (GRADIENTFLOW/GradientFlow.cpp:168-188) */
double or_flow_calibration(const or_params *p, double *U, double *E, int n, double dtau) {
    double ac = p->a[0]*p->a[1]*p->a[2];
    double DcTau = dtau*pow(ac, 2.0/3.0)/(p->aScale*p->aScale), d = 0.0;
    for (int k = 0; k < n; k++) {
        or_flow_force(p, U, E);
        d += 0.5*DcTau*or_ncsdot(p, U, E);
        or_flow_links(p, U, E, dtau);
        d += 0.5*DcTau*or_ncsdot(p, U, E);
    }
    return d;
}
void or_block_links(const or_params *p, const double *U, double *Uc) {
    DISPATCH(p, su2l_block_links(p, U, Uc), su3l_block_links(p, U, Uc));
}
void or_hard_scales(const or_params *p, const double *U, double out[3]) {
    DISPATCH(p, su2l_hard_scales(p, U, out), su3l_hard_scales(p, U, out));
}
void or_bulk(const or_params *p, const double *U, const double *E, double out[6]) {
    DISPATCH(p, su2l_bulk(p, U, E, out), su3l_bulk(p, U, E, out));
}
void or_gauss(const or_params *p, const double *U, const double *E, double *smax, double *rms) {
    DISPATCH(p, su2l_gauss(p, U, E, smax, rms), su3l_gauss(p, U, E, smax, rms));
}
void or_gauss_field(const or_params *p, const double *U, const double *E, double *G) {
    DISPATCH(p, su2l_gauss_field(p, U, E, G), su3l_gauss_field(p, U, E, G));
}
void or_unitarity(const or_params *p, const double *U, double *umax, double *rms) {
    DISPATCH(p, su2l_unitarity(p, U, umax, rms), su3l_unitarity(p, U, umax, rms));
}
int or_gauss_restore(const or_params *p, const double *U, double *E, int maxiter, double tol, double *fm) {
    int r = 0; DISPATCH(p, r = su2l_gauss_restore(p, U, E, maxiter, tol, fm), r = su3l_gauss_restore(p, U, E, maxiter, tol, fm)); return r;
}
void or_geodesic(const or_params *p, const double *Uo, const double *Un, double c, double *Um, double *Em) {
    DISPATCH(p, su2l_geodesic(p, Uo, Un, c, Um, Em), su3l_geodesic(p, Uo, Un, c, Um, Em));
}
/* CoolingMethod::GetDeltaNCs(1): Simpson (1,4,1)/6 over the geodesic
   between two cooled configurations (TOPOLOGY/CoolingMethod.cpp:80-121) */
double or_delta_ncs_cooling(const or_params *p, const double *Uo, const double *Un, double *Um, double *Em) {
    double v = 0.0, w = 0.0;
    or_geodesic(p, Uo, Un, 0.5, Um, Em);
    v += 4.0*or_ncsdot(p, Um, Em); w += 4.0;
    v += or_ncsdot(p, Uo, Em); w += 1.0;
    v += or_ncsdot(p, Un, Em); w += 1.0;
    return v/w;
}
double or_vacuum_ncs(const or_params *p, const double *U) {
    double r = 0.0; DISPATCH(p, r = su2l_vacuum_ncs(p, U), r = su3l_vacuum_ncs(p, U)); return r;
}
void or_reunitarize_field(const or_params *p, double *U) {
    DISPATCH(p, su2l_reunitarize_field(p, U), su3l_reunitarize_field(p, U));
}

/* ThermalDynamics::Thermalize(cycles, max_time) (DYNAMICS/ThermalDynamics.cpp:260-316):
   per cycle, ReassignElectricFields and GaussLawRestoration::Restore (:205-226),
   then leapfrog steps while tSteps*dTau < MaxTime with a Bulk row
   {Time() + i MaxTime, T00, TXX, TYY, TZZ, ELECTRIC, MAGNETIC} every 20 steps
   (:274-292; BulkObservables.cpp:17-52, Minkowski); Reset (:85-91) after each
   cycle.  Returns the number of rows (at most log_rows are stored). */
int or_thermalize(const or_params *p, double *U, double *E, int cycles, double max_time, double dt, double beta,
                  uint64_t seed, int maxiter, double tol, double *log, int log_rows) {
    double cb[3];
    for (int mu = 0; mu < 3; mu++) {
        const double x = p->a[mu]*p->aScale*p->aScale/(p->a[0]*p->a[1]*p->a[2]);
        cb[mu] = x*x;
    }
    int rows = 0;
    for (int i = 0; i < cycles; i++) {
        double fm;
        or_thermal_efield(p, seed, (uint32_t)i, beta, E);
        or_gauss_restore(p, U, E, maxiter, tol, &fm);
        long t = 0;
        while ((double)t*dt < max_time) {
            or_leapfrog(p, U, E, 1, dt);
            t++;
            if (t % 20 == 0) {
                double b[6];
                or_bulk(p, U, E, b);
                if (rows < log_rows) {
                    double *w = log + 7*(size_t)rows;
                    const double E0 = cb[0]*b[0], E1 = cb[1]*b[1], E2 = cb[2]*b[2];
                    const double B0 = cb[0]*b[3], B1 = cb[1]*b[4], B2 = cb[2]*b[5];
                    w[0] = (double)t*dt + i*max_time;
                    w[1] = 0.5*(E0 + E1 + E2 + B0 + B1 + B2);
                    w[2] = 0.5*((E1 + E2 + B1 + B2) - (E0 + B0));
                    w[3] = 0.5*((E0 + E2 + B0 + B2) - (E1 + B1));
                    w[4] = 0.5*((E0 + E1 + B0 + B1) - (E2 + B2));
                    w[5] = 0.5*(E0 + E1 + E2);
                    w[6] = 0.5*(B0 + B1 + B2);
                }
                rows++;
            }
        }
    }
    return rows;
}
