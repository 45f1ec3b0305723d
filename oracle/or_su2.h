/*
 * SU(2) primitives of the oracle.  Test infrastructure only.
 *
 * A link is a quaternion (u0,u1,u2,u3) with U = u3*1 + i u_a sigma_a
 * (SUNc/SU2/GroupOperations.cpp:4-9, matrix form :213-225).  Real linear
 * combinations of links (clover sums) stay in that span, so every routine
 * below is valid for them as well.  Generators t^a = sigma_a/2
 * (SUNc/SU2/AlgebraOperations.cpp:4-6).
 */
#ifndef OR_SU2_H
#define OR_SU2_H
#include <math.h>

#define SU2_SU 4
#define SU2_NE 3

/* quaternion product: (a3 + i a.s)(b3 + i b.s) = (a3 b3 - a.b) + i(a3 b + b3 a - a x b).s
   (SUNc/SU2/GroupOperations.cpp:103-110) */
static inline void su2_mul(const double *A, const double *B, double *C) {
    double c0 = A[3]*B[0] + B[3]*A[0] - (A[1]*B[2] - A[2]*B[1]);
    double c1 = A[3]*B[1] + B[3]*A[1] - (A[2]*B[0] - A[0]*B[2]);
    double c2 = A[3]*B[2] + B[3]*A[2] - (A[0]*B[1] - A[1]*B[0]);
    double c3 = A[3]*B[3] - A[0]*B[0] - A[1]*B[1] - A[2]*B[2];
    C[0] = c0; C[1] = c1; C[2] = c2; C[3] = c3;
}
static inline void su2_dag(const double *A, double *D) {
    D[0] = -A[0]; D[1] = -A[1]; D[2] = -A[2]; D[3] = A[3];
}
/* A B^+  (:113-120) */
static inline void su2_mul_nd(const double *A, const double *B, double *C) {
    double Bd[4]; su2_dag(B, Bd); su2_mul(A, Bd, C);
}
/* A^+ B  (:122-129) */
static inline void su2_mul_dn(const double *A, const double *B, double *C) {
    double Ad[4]; su2_dag(A, Ad); su2_mul(Ad, B, C);
}
/* A^+ B^+ (:131-138) */
static inline void su2_mul_dd(const double *A, const double *B, double *C) {
    double Ad[4], Bd[4]; su2_dag(A, Ad); su2_dag(B, Bd); su2_mul(Ad, Bd, C);
}
/* c ReTr(i t^a M) = -c m_a  (:78-92) */
static inline void su2_proj(double c, const double *M, double *out) {
    out[0] = -c*M[0]; out[1] = -c*M[1]; out[2] = -c*M[2];
}
/* ReTr(1 - M) = 2(1 - m3)  (:70-72) */
static inline double su2_retr_1m(const double *M) { return 2.0*(1.0 - M[3]); }

/* exp(i c alpha^a sigma_a/2) in closed form (SUNc/SU2/AlgebraOperations.cpp:37-55) */
static inline void su2_expi(double c, const double *al, double *U) {
    double d = sqrt((c*al[0])*(c*al[0]) + (c*al[1])*(c*al[1]) + (c*al[2])*(c*al[2]));
    if (d != 0.0) {
        double h = 0.5*d, m = c*sin(h)/d;
        U[0] = m*al[0]; U[1] = m*al[1]; U[2] = m*al[2]; U[3] = cos(h);
    } else { U[0] = U[1] = U[2] = 0.0; U[3] = 1.0; }
}
/* alpha with exp(i c alpha^a sigma_a/2) = U, principal branch via atan2
   (SUNc/SU2/AlgebraOperations.cpp:59-74) */
static inline void su2_logi(double c, const double *U, double *al) {
    double d = sqrt(U[0]*U[0] + U[1]*U[1] + U[2]*U[2]);
    if (d != 0.0) {
        double m = 2.0*atan2(d, U[3])/(c*d);
        al[0] = m*U[0]; al[1] = m*U[1]; al[2] = m*U[2];
    } else { al[0] = al[1] = al[2] = 0.0; }
}
/* (V E V^+)^a = (w^2-|v|^2) E + 2 v (v.E) + 2 w (E x v), V = w + i v.sigma
   (SUNc/SU2/AlgebraOperations.cpp:77-83, derived from U^adj_ab = 2tr[t^a U t^b U^+]) */
static inline void su2_adj_act(const double *V, const double *E, double *o) {
    double w = V[3], vv = V[0]*V[0] + V[1]*V[1] + V[2]*V[2];
    double ve = V[0]*E[0] + V[1]*E[1] + V[2]*E[2], s = w*w - vv;
    double x0 = E[1]*V[2] - E[2]*V[1], x1 = E[2]*V[0] - E[0]*V[2], x2 = E[0]*V[1] - E[1]*V[0];
    o[0] = s*E[0] + 2.0*V[0]*ve + 2.0*w*x0;
    o[1] = s*E[1] + 2.0*V[1]*ve + 2.0*w*x1;
    o[2] = s*E[2] + 2.0*V[2]*ve + 2.0*w*x2;
}
/* (V^+ E V)^a  (:86-91) */
static inline void su2_inv_adj_act(const double *V, const double *E, double *o) {
    double Vd[4]; su2_dag(V, Vd); su2_adj_act(Vd, E, o);
}
/* ||U U^+ - 1|| over the 4 quaternion components (:172-184) */
static inline double su2_unitarity_norm(const double *U) {
    double C[4]; su2_mul_nd(U, U, C);
    double n = C[0]*C[0] + C[1]*C[1] + C[2]*C[2] + (C[3]-1.0)*(C[3]-1.0);
    return sqrt(n);
}
static inline void su2_reunitarize(double *U) {
    double n = sqrt(U[0]*U[0] + U[1]*U[1] + U[2]*U[2] + U[3]*U[3]);
    U[0] /= n; U[1] /= n; U[2] /= n; U[3] /= n;
}
static inline void su2_unit(double *U) { U[0] = U[1] = U[2] = 0.0; U[3] = 1.0; }
/* structure constants f^{abc} = eps_{abc} (SUNc/SU2/AlgebraOperations.cpp:23-25) */
static inline double su2_f(int a, int b, int c) { return ((a-b)*(a-c)*(c-b))/2.0; }

#endif
