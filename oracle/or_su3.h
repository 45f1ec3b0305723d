/*
 * SU(3) primitives of the oracle.  Test infrastructure only.
 *
 * A link is 9 complex numbers stored as 18 doubles (re,im), column-major
 * element (i,j) at complex index i+3j (SUNc/SU3/GroupOperations.cpp:4-14,28-30).
 * Generators t^a = lambda_a/2 in Gell-Mann order
 * (SUNc/SU3/AlgebraOperations.cpp:4-7, matrix form :89-103).
 *
 * exp: hermitian eigen-decomposition U = sum_j v_j e^{i c w_j} v_j^+ as in
 *      SUNc/SU3/AlgebraOperations.cpp:141-182.  Eigenvalues from the
 *      trigonometric Cardano solution, eigenvectors from cross products of the
 *      rows of (A - w) (the hybrid analytic method of Kopp 2008, rewritten from
 *      its published description), and a complex Jacobi sweep whenever two
 *      eigenvalues are closer than 1e-3 of the spectral scale.
 * log: principal eigen-phases as in SUNc/SU3/AlgebraOperations.cpp:186-233
 *      (which calls LAPACK zgeev); here the eigenvectors come from a Jacobi
 *      diagonalisation of the Cayley transform i(1-wW)(1+wW)^-1 of a
 *      phase-shifted W, which shares W's eigenvectors.
 */
#ifndef OR_SU3_H
#define OR_SU3_H
#include <math.h>

#define SU3_SU 18
#define SU3_NE 8

typedef struct { double re, im; } ocd;

#define OSQRT3 1.7320508075688772935274463415058723669428052538103806280558

static inline ocd oc(double r, double i) { ocd z; z.re = r; z.im = i; return z; }
static inline ocd oadd(ocd a, ocd b) { return oc(a.re + b.re, a.im + b.im); }
static inline ocd osub(ocd a, ocd b) { return oc(a.re - b.re, a.im - b.im); }
static inline ocd omul(ocd a, ocd b) { return oc(a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re); }
static inline ocd ocj(ocd a) { return oc(a.re, -a.im); }
static inline ocd oscale(double s, ocd a) { return oc(s*a.re, s*a.im); }
static inline double oabs2(ocd a) { return a.re*a.re + a.im*a.im; }

#define M3(M,i,j) ((M)[(i)+3*(j)])

/* C = A B  (SUNc/SU3/GroupOperations.cpp:148-162) */
static inline void su3_mul(const double *Ad, const double *Bd, double *Cd) {
    const ocd *A = (const ocd*)Ad, *B = (const ocd*)Bd; ocd C[9];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
        ocd s = oc(0, 0);
        for (int k = 0; k < 3; k++) s = oadd(s, omul(M3(A,i,k), M3(B,k,j)));
        M3(C,i,j) = s;
    }
    for (int s = 0; s < 9; s++) ((ocd*)Cd)[s] = C[s];
}
/* C = A B^+  (:165-179) */
static inline void su3_mul_nd(const double *Ad, const double *Bd, double *Cd) {
    const ocd *A = (const ocd*)Ad, *B = (const ocd*)Bd; ocd C[9];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
        ocd s = oc(0, 0);
        for (int k = 0; k < 3; k++) s = oadd(s, omul(M3(A,i,k), ocj(M3(B,j,k))));
        M3(C,i,j) = s;
    }
    for (int s = 0; s < 9; s++) ((ocd*)Cd)[s] = C[s];
}
/* C = A^+ B  (:181-195) */
static inline void su3_mul_dn(const double *Ad, const double *Bd, double *Cd) {
    const ocd *A = (const ocd*)Ad, *B = (const ocd*)Bd; ocd C[9];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
        ocd s = oc(0, 0);
        for (int k = 0; k < 3; k++) s = oadd(s, omul(ocj(M3(A,k,i)), M3(B,k,j)));
        M3(C,i,j) = s;
    }
    for (int s = 0; s < 9; s++) ((ocd*)Cd)[s] = C[s];
}
/* C = A^+ B^+  (:197-211) */
static inline void su3_mul_dd(const double *Ad, const double *Bd, double *Cd) {
    const ocd *A = (const ocd*)Ad, *B = (const ocd*)Bd; ocd C[9];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
        ocd s = oc(0, 0);
        for (int k = 0; k < 3; k++) s = oadd(s, omul(ocj(M3(A,k,i)), ocj(M3(B,j,k))));
        M3(C,i,j) = s;
    }
    for (int s = 0; s < 9; s++) ((ocd*)Cd)[s] = C[s];
}
static inline void su3_dag(const double *Ad, double *Dd) {
    const ocd *A = (const ocd*)Ad; ocd D[9];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) M3(D,i,j) = ocj(M3(A,j,i));
    for (int s = 0; s < 9; s++) ((ocd*)Dd)[s] = D[s];
}
/* c ReTr(i t^a M) for the 8 Gell-Mann generators (:108-130):
   tr(i lambda_a M)/2 worked out entry by entry. */
static inline void su3_proj(double c, const double *Md, double *o) {
    const ocd *M = (const ocd*)Md;
    double h = -0.5*c;
    o[0] = h*(M3(M,0,1).im + M3(M,1,0).im);
    o[1] = h*(M3(M,0,1).re - M3(M,1,0).re);
    o[2] = h*(M3(M,0,0).im - M3(M,1,1).im);
    o[3] = h*(M3(M,0,2).im + M3(M,2,0).im);
    o[4] = h*(M3(M,0,2).re - M3(M,2,0).re);
    o[5] = h*(M3(M,1,2).im + M3(M,2,1).im);
    o[6] = h*(M3(M,1,2).re - M3(M,2,1).re);
    o[7] = h*(M3(M,0,0).im + M3(M,1,1).im - 2.0*M3(M,2,2).im)/OSQRT3;
}
/* ReTr(1 - M) (:86-88) */
static inline double su3_retr_1m(const double *Md) {
    const ocd *M = (const ocd*)Md;
    return 3.0 - M3(M,0,0).re - M3(M,1,1).re - M3(M,2,2).re;
}
/* A = alpha^a t^a as a hermitian matrix (:89-103) */
static inline void su3_alg_matrix(const double *al, ocd *A) {
    M3(A,0,0) = oc(0.5*(al[2] + al[7]/OSQRT3), 0);
    M3(A,1,1) = oc(-0.5*(al[2] - al[7]/OSQRT3), 0);
    M3(A,2,2) = oc(-al[7]/OSQRT3, 0);
    M3(A,1,0) = oc(0.5*al[0], 0.5*al[1]);
    M3(A,2,0) = oc(0.5*al[3], 0.5*al[4]);
    M3(A,2,1) = oc(0.5*al[5], 0.5*al[6]);
    M3(A,0,1) = ocj(M3(A,1,0));
    M3(A,0,2) = ocj(M3(A,2,0));
    M3(A,1,2) = ocj(M3(A,2,1));
}
/* alpha^a = 2 tr(t^a X) = tr(lambda_a X) for hermitian X */
static inline void su3_herm_alg(const ocd *X, double *o) {
    o[0] = M3(X,0,1).re + M3(X,1,0).re;
    o[1] = M3(X,1,0).im - M3(X,0,1).im;
    o[2] = M3(X,0,0).re - M3(X,1,1).re;
    o[3] = M3(X,0,2).re + M3(X,2,0).re;
    o[4] = M3(X,2,0).im - M3(X,0,2).im;
    o[5] = M3(X,1,2).re + M3(X,2,1).re;
    o[6] = M3(X,2,1).im - M3(X,1,2).im;
    o[7] = (M3(X,0,0).re + M3(X,1,1).re - 2.0*M3(X,2,2).re)/OSQRT3;
}
/* (V E V^+)^a with E = E^b t^b; equals U^adj_ab E^b with
   U^adj_ab = 2 tr[t^a V t^b V^+] (SUNc/SU3/AlgebraOperations.cpp:257-275) */
static inline void su3_adj_act(const double *V, const double *E, double *o) {
    ocd A[9], T[9];
    su3_alg_matrix(E, A);
    su3_mul(V, (const double*)A, (double*)T);
    su3_mul_nd((const double*)T, V, (double*)A);
    su3_herm_alg(A, o);
}
/* (V^+ E V)^a (:278-295) */
static inline void su3_inv_adj_act(const double *V, const double *E, double *o) {
    ocd A[9], T[9];
    su3_alg_matrix(E, A);
    su3_mul_dn(V, (const double*)A, (double*)T);
    su3_mul((const double*)T, V, (double*)A);
    su3_herm_alg(A, o);
}
/* Frobenius ||U U^+ - 1|| (:351-363) */
static inline double su3_unitarity_norm(const double *U) {
    ocd C[9]; su3_mul_nd(U, U, (double*)C);
    double n = 0.0;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
        ocd d = M3(C,i,j); if (i == j) d.re -= 1.0; n += oabs2(d);
    }
    return sqrt(n);
}
static inline void su3_unit(double *U) {
    for (int s = 0; s < 18; s++) U[s] = 0.0;
    U[0] = U[8] = U[16] = 1.0;
}

/* ---------- hermitian 3x3 eigen-decomposition ---------- */

/* complex Jacobi: A (hermitian, overwritten) -> diag; columns of V = eigenvectors */
static inline void su3_jacobi(ocd *A, ocd *V, double *w) {
    for (int s = 0; s < 9; s++) V[s] = oc(0, 0);
    M3(V,0,0) = M3(V,1,1) = M3(V,2,2) = oc(1, 0);
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = oabs2(M3(A,0,1)) + oabs2(M3(A,0,2)) + oabs2(M3(A,1,2));
        double dia = M3(A,0,0).re*M3(A,0,0).re + M3(A,1,1).re*M3(A,1,1).re + M3(A,2,2).re*M3(A,2,2).re;
        if (off == 0.0 || off < 1e-36*dia) break;
        for (int p = 0; p < 2; p++) for (int q = p + 1; q < 3; q++) {
            ocd apq = M3(A,p,q);
            double g = sqrt(oabs2(apq));
            if (g == 0.0) continue;
            ocd e = oscale(1.0/g, apq);                  /* unit phase of A_pq */
            double th = (M3(A,q,q).re - M3(A,p,p).re)/(2.0*g);
            double t = (th >= 0 ? 1.0 : -1.0)/(fabs(th) + sqrt(th*th + 1.0));
            double c = 1.0/sqrt(t*t + 1.0), sn = t*c;
            /* R = D J: D = diag(1, conj e) on (p,q), J real rotation */
            ocd R[9];
            for (int k = 0; k < 9; k++) R[k] = oc(0, 0);
            for (int k = 0; k < 3; k++) M3(R,k,k) = oc(1, 0);
            M3(R,p,p) = oc(c, 0);
            M3(R,p,q) = oc(sn, 0);
            M3(R,q,p) = oscale(-sn, ocj(e));
            M3(R,q,q) = oscale(c, ocj(e));
            ocd T[9];
            su3_mul((const double*)A, (const double*)R, (double*)T);
            su3_mul_dn((const double*)R, (const double*)T, (double*)A);
            su3_mul((const double*)V, (const double*)R, (double*)T);
            for (int k = 0; k < 9; k++) V[k] = T[k];
            M3(A,p,q) = M3(A,q,p) = oc(0, 0);
            for (int k = 0; k < 3; k++) M3(A,k,k).im = 0.0;
        }
    }
    for (int k = 0; k < 3; k++) w[k] = M3(A,k,k).re;
}

/* complex cross product a x b (bilinear) */
static inline void ocross(const ocd *a, const ocd *b, ocd *c) {
    c[0] = osub(omul(a[1], b[2]), omul(a[2], b[1]));
    c[1] = osub(omul(a[2], b[0]), omul(a[0], b[2]));
    c[2] = osub(omul(a[0], b[1]), omul(a[1], b[0]));
}

/* eigenvector of hermitian A for eigenvalue lam from the best-conditioned
   cross product of two rows of (A - lam); returns squared norm before
   normalisation relative to the row norms */
static inline double su3_eigvec_cross(const ocd *A, double lam, ocd *v) {
    ocd r[3][3];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
        r[i][j] = M3(A,i,j); if (i == j) r[i][j].re -= lam;
    }
    ocd c[3][3]; double n[3], best = -1.0; int ib = 0;
    ocross(r[0], r[1], c[0]); ocross(r[0], r[2], c[1]); ocross(r[1], r[2], c[2]);
    for (int k = 0; k < 3; k++) {
        n[k] = oabs2(c[k][0]) + oabs2(c[k][1]) + oabs2(c[k][2]);
        if (n[k] > best) { best = n[k]; ib = k; }
    }
    double s = 1.0/sqrt(best);
    for (int k = 0; k < 3; k++) v[k] = oscale(s, c[ib][k]);
    return best;
}

/* eigen-decomposition: V columns are eigenvectors, w eigenvalues */
static inline void su3_heig(const ocd *Ain, ocd *V, double *w) {
    /* traceless hermitian: lambda^3 - p lambda - q = 0 */
    double p = 0.0;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) p += oabs2(M3(Ain,i,j));
    p *= 0.5;
    double tr = (M3(Ain,0,0).re + M3(Ain,1,1).re + M3(Ain,2,2).re)/3.0;
    if (p == 0.0 && tr == 0.0) {
        for (int s = 0; s < 9; s++) V[s] = oc(0, 0);
        M3(V,0,0) = M3(V,1,1) = M3(V,2,2) = oc(1, 0);
        w[0] = w[1] = w[2] = 0.0; return;
    }
    ocd A[9];
    for (int s = 0; s < 9; s++) A[s] = Ain[s];
    for (int k = 0; k < 3; k++) M3(A,k,k).re -= tr;     /* remove trace (0 for algebra input) */
    p = 0.0;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) p += oabs2(M3(A,i,j));
    p *= 0.5;
    /* q = det A */
    ocd d0 = osub(omul(M3(A,1,1), M3(A,2,2)), omul(M3(A,1,2), M3(A,2,1)));
    ocd d1 = osub(omul(M3(A,1,0), M3(A,2,2)), omul(M3(A,1,2), M3(A,2,0)));
    ocd d2 = osub(omul(M3(A,1,0), M3(A,2,1)), omul(M3(A,1,1), M3(A,2,0)));
    double q = omul(M3(A,0,0), d0).re - omul(M3(A,0,1), d1).re + omul(M3(A,0,2), d2).re;
    int ok = p > 0.0;
    double lam[3];
    if (ok) {
        double r = sqrt(p/3.0);
        double disc = 4.0*p*p*p/27.0 - q*q;
        double phi = atan2(sqrt(disc > 0.0 ? disc : 0.0), q)/3.0;
        double scale = sqrt(p);
        for (int k = 0; k < 3; k++) {
            double l = 2.0*r*cos(phi - 2.0*M_PI*k/3.0);
            /* one Newton step, kept only as a refinement: at a (near) double
               root the derivative 3 l^2 - p vanishes and the step would throw
               l far off (and past the degeneracy test below) */
            double dp = 3.0*l*l - p;
            if (dp != 0.0) {
                double st = (l*l*l - p*l - q)/dp;
                if (fabs(st) < 1e-6*scale) l -= st;
            }
            lam[k] = l;
        }
        double g01 = fabs(lam[0]-lam[1]), g02 = fabs(lam[0]-lam[2]), g12 = fabs(lam[1]-lam[2]);
        double gmin = g01 < g02 ? g01 : g02; if (g12 < gmin) gmin = g12;
        ok = gmin >= 1e-3*scale;
    }
    if (ok) {
        ocd v0[3], v1[3], v2[3];
        su3_eigvec_cross(A, lam[0], v0);
        su3_eigvec_cross(A, lam[1], v1);
        /* orthogonalise v1 against v0, then v2 = conj(v0 x v1) */
        ocd ip = oc(0, 0);
        for (int k = 0; k < 3; k++) ip = oadd(ip, omul(ocj(v0[k]), v1[k]));
        double nn = 0.0;
        for (int k = 0; k < 3; k++) { v1[k] = osub(v1[k], omul(ip, v0[k])); nn += oabs2(v1[k]); }
        nn = 1.0/sqrt(nn);
        for (int k = 0; k < 3; k++) v1[k] = oscale(nn, v1[k]);
        ocross(v0, v1, v2);
        for (int k = 0; k < 3; k++) v2[k] = ocj(v2[k]);
        for (int i = 0; i < 3; i++) {
            M3(V,i,0) = v0[i]; M3(V,i,1) = v1[i]; M3(V,i,2) = v2[i];
        }
        w[0] = lam[0] + tr; w[1] = lam[1] + tr; w[2] = lam[2] + tr;
        return;
    }
    su3_jacobi(A, V, w);
    w[0] += tr; w[1] += tr; w[2] += tr;
}

/* exp(i c alpha^a t^a) (SUNc/SU3/AlgebraOperations.cpp:141-182) */
static inline void su3_expi(double c, const double *al, double *Ud) {
    ocd A[9], V[9]; double w[3];
    su3_alg_matrix(al, A);
    su3_heig(A, V, w);
    ocd *U = (ocd*)Ud;
    for (int s = 0; s < 9; s++) U[s] = oc(0, 0);
    for (int j = 0; j < 3; j++) {
        ocd e = oc(cos(c*w[j]), sin(c*w[j]));
        for (int i = 0; i < 3; i++) for (int k = 0; k < 3; k++)
            M3(U,i,k) = oadd(M3(U,i,k), omul(omul(M3(V,i,j), e), ocj(M3(V,k,j))));
    }
}

/* 3x3 complex inverse via the adjugate */
static inline int su3_inverse(const ocd *M, ocd *I) {
    ocd c00 = osub(omul(M3(M,1,1),M3(M,2,2)), omul(M3(M,1,2),M3(M,2,1)));
    ocd c01 = osub(omul(M3(M,1,2),M3(M,2,0)), omul(M3(M,1,0),M3(M,2,2)));
    ocd c02 = osub(omul(M3(M,1,0),M3(M,2,1)), omul(M3(M,1,1),M3(M,2,0)));
    ocd det = oadd(oadd(omul(M3(M,0,0),c00), omul(M3(M,0,1),c01)), omul(M3(M,0,2),c02));
    double dd = oabs2(det);
    if (dd == 0.0) return -1;
    ocd id = oscale(1.0/dd, ocj(det));
    ocd c10 = osub(omul(M3(M,0,2),M3(M,2,1)), omul(M3(M,0,1),M3(M,2,2)));
    ocd c11 = osub(omul(M3(M,0,0),M3(M,2,2)), omul(M3(M,0,2),M3(M,2,0)));
    ocd c12 = osub(omul(M3(M,0,1),M3(M,2,0)), omul(M3(M,0,0),M3(M,2,1)));
    ocd c20 = osub(omul(M3(M,0,1),M3(M,1,2)), omul(M3(M,0,2),M3(M,1,1)));
    ocd c21 = osub(omul(M3(M,0,2),M3(M,1,0)), omul(M3(M,0,0),M3(M,1,2)));
    ocd c22 = osub(omul(M3(M,0,0),M3(M,1,1)), omul(M3(M,0,1),M3(M,1,0)));
    /* inverse = adj / det, adj_ij = cofactor_ji */
    M3(I,0,0) = omul(c00,id); M3(I,0,1) = omul(c10,id); M3(I,0,2) = omul(c20,id);
    M3(I,1,0) = omul(c01,id); M3(I,1,1) = omul(c11,id); M3(I,1,2) = omul(c21,id);
    M3(I,2,0) = omul(c02,id); M3(I,2,1) = omul(c12,id); M3(I,2,2) = omul(c22,id);
    return 0;
}

/* alpha with exp(i c alpha^a t^a) = U, principal eigen-phases in (-pi, pi],
   projected onto the generators (SUNc/SU3/AlgebraOperations.cpp:186-233).
   The reference returns c*h for SU(3) and h/c for SU(2); both agree at
   c = +-1, the only values the hot path uses (SUNc/AdvancedOperations.cpp:432-442). */
static inline void su3_logi(double c, const double *Ud, double *al) {
    const ocd *W = (const ocd*)Ud;
    /* pick a global phase shift e^{-i s} keeping all eigenvalues away from -1 */
    static const double shifts[4] = {0.0, 1.5707963267948966, -1.5707963267948966, 3.141592653589793};
    double best = -1.0; int sb = 0;
    for (int k = 0; k < 4; k++) {
        ocd ph = oc(cos(shifts[k]), -sin(shifts[k])), M[9];
        for (int s = 0; s < 9; s++) M[s] = omul(ph, W[s]);
        for (int i = 0; i < 3; i++) M3(M,i,i).re += 1.0;
        ocd c00 = osub(omul(M3(M,1,1),M3(M,2,2)), omul(M3(M,1,2),M3(M,2,1)));
        ocd c01 = osub(omul(M3(M,1,2),M3(M,2,0)), omul(M3(M,1,0),M3(M,2,2)));
        ocd c02 = osub(omul(M3(M,1,0),M3(M,2,1)), omul(M3(M,1,1),M3(M,2,0)));
        ocd det = oadd(oadd(omul(M3(M,0,0),c00), omul(M3(M,0,1),c01)), omul(M3(M,0,2),c02));
        double ad = oabs2(det);
        if (ad > best) { best = ad; sb = k; }
        if (k == 0 && ad > 0.25) break;        /* all |1+e^{i phi}| comfortably > 0 */
    }
    double sft = shifts[sb];
    ocd ph = oc(cos(sft), -sin(sft)), Wp[9], P[9], Mi[9], Cm[9];
    for (int s = 0; s < 9; s++) Wp[s] = omul(ph, W[s]);
    for (int s = 0; s < 9; s++) { P[s] = Wp[s]; Mi[s] = oscale(-1.0, Wp[s]); }
    for (int i = 0; i < 3; i++) { M3(P,i,i).re += 1.0; M3(Mi,i,i).re += 1.0; }
    ocd Pinv[9];
    su3_inverse(P, Pinv);
    su3_mul((const double*)Mi, (const double*)Pinv, (double*)Cm);   /* (1-W')(1+W')^-1 */
    /* C = i (1-W')(1+W')^-1 is hermitian; symmetrise */
    ocd C[9];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
        ocd a = omul(oc(0, 1), M3(Cm,i,j)), b = ocj(omul(oc(0, 1), M3(Cm,j,i)));
        M3(C,i,j) = oscale(0.5, oadd(a, b));
    }
    ocd V[9]; double w[3];
    su3_jacobi(C, V, w);
    ocd H[9];
    for (int s = 0; s < 9; s++) H[s] = oc(0, 0);
    for (int j = 0; j < 3; j++) {
        double phi = 2.0*atan(w[j]) + sft;          /* eigen-phase of W */
        while (phi > M_PI) phi -= 2.0*M_PI;
        while (phi <= -M_PI) phi += 2.0*M_PI;
        for (int i = 0; i < 3; i++) for (int k = 0; k < 3; k++)
            M3(H,i,k) = oadd(M3(H,i,k), oscale(phi, omul(M3(V,i,j), ocj(M3(V,k,j)))));
    }
    su3_herm_alg(H, al);
    for (int a = 0; a < 8; a++) al[a] /= c;
}

/* Gram-Schmidt on rows 0,1 and row2 = conj(row0 x row1) */
static inline void su3_reunitarize(double *Ud) {
    ocd *U = (ocd*)Ud; ocd r0[3], r1[3], r2[3];
    for (int j = 0; j < 3; j++) { r0[j] = M3(U,0,j); r1[j] = M3(U,1,j); }
    double n = 0.0;
    for (int j = 0; j < 3; j++) n += oabs2(r0[j]);
    n = 1.0/sqrt(n);
    for (int j = 0; j < 3; j++) r0[j] = oscale(n, r0[j]);
    ocd ip = oc(0, 0);
    for (int j = 0; j < 3; j++) ip = oadd(ip, omul(ocj(r0[j]), r1[j]));
    n = 0.0;
    for (int j = 0; j < 3; j++) { r1[j] = osub(r1[j], omul(ip, r0[j])); n += oabs2(r1[j]); }
    n = 1.0/sqrt(n);
    for (int j = 0; j < 3; j++) r1[j] = oscale(n, r1[j]);
    ocross(r0, r1, r2);
    for (int j = 0; j < 3; j++) {
        M3(U,0,j) = r0[j]; M3(U,1,j) = r1[j]; M3(U,2,j) = ocj(r2[j]);
    }
}

/* structure constants of su(3), 0-based Gell-Mann indices
   (SUNc/SU3/AlgebraOperations.cpp:27-80) */
static inline double su3_f(int a, int b, int c) {
    static double F[8][8][8]; static int init = 0;
    if (!init) {
        for (int i = 0; i < 512; i++) ((double*)F)[i] = 0.0;
        const int T[9][3] = {{0,1,2},{0,3,6},{0,4,5},{1,3,5},{1,4,6},{2,3,4},{2,5,6},{3,4,7},{5,6,7}};
        const double v[9] = {1.0, 0.5, -0.5, 0.5, 0.5, 0.5, -0.5, 0.8660254037844386, 0.8660254037844386};
        for (int t = 0; t < 9; t++) {
            int i = T[t][0], j = T[t][1], k = T[t][2]; double x = v[t];
            F[i][j][k] = x; F[j][k][i] = x; F[k][i][j] = x;
            F[j][i][k] = -x; F[i][k][j] = -x; F[k][j][i] = -x;
        }
        init = 1;
    }
    return F[a][b][c];
}

#endif
