// Device-side SU(2)/SU(3) algebra for the gfx950 kernels.
//
// Storage conventions are those of the reference's GaugeLinks/ElectricFields:
//   SU(2): 4 reals (u0,u1,u2,u3), U = u3 + i u_a sigma_a   (SUNc/SU2/GroupOperations.cpp:4-9)
//   SU(3): 9 complex, column-major (i,j) -> i+3j, stored re,im (SUNc/SU3/GroupOperations.cpp:4-14)
//   algebra: t^a = sigma_a/2 or lambda_a/2 (Gell-Mann order)
// Every routine is templated on the arithmetic type T (double for the FP64
// form, float for the FP32 form).  Reunitarisation runs in double (Reunit<T>)
// in both forms.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace tcym {

template <typename T> using Reunit = double;
// 1/sqrt(x): correctly rounded division and square root in double; the
// hardware reciprocal square root refined by one Newton step in float
__device__ __forceinline__ double rsqrt_of(double x) { return 1.0/__builtin_sqrt(x); }
__device__ __forceinline__ float  rsqrt_of(float x) {
    const float r = __builtin_amdgcn_rsqf(x);
    return r*(1.5f - 0.5f*x*r*r);   // one Newton step on the hardware estimate
}
// 1/sqrt(x) of a squared row norm in the reunitarisation of a link stored in
// T.  FP64 links: rsqrt_of.  FP32 links reunitarised in double: the hardware
// double estimate refined by two Newton steps (9 instructions instead of the
// 31 of a correctly rounded square root and division), within a few double
// ulp of 1/sqrt(x); the FP32 link it rounds to differs from the one of the
// exact value only where that lies within those ulp of a float boundary.
template <typename T>
__device__ __forceinline__ double rsqrt_reunit(double x) {
    if constexpr (sizeof(T) == 8) {
        return rsqrt_of(x);
    } else {
        const double h = 0.5*x;
        double r = __builtin_amdgcn_rsq(x);
        r = r*__builtin_fma(-h*r, r, 1.5);
        return r*__builtin_fma(-h*r, r, 1.5);
    }
}
template <typename T>
__device__ __forceinline__ float rsqrt_reunit(float x) { return rsqrt_of(x); }

// 1/k!, k = 0..16
__device__ constexpr double kInvFact[17] = {
    1.0, 1.0, 0.5, 1.0/6.0, 1.0/24.0, 1.0/120.0, 1.0/720.0, 1.0/5040.0, 1.0/40320.0,
    1.0/362880.0, 1.0/3628800.0, 1.0/39916800.0, 1.0/479001600.0, 1.0/6227020800.0,
    1.0/87178291200.0, 1.0/1307674368000.0, 1.0/20922789888000.0};

template <typename T> struct Lit;
template <> struct Lit<double> { static constexpr int exp_terms = 15; };
template <> struct Lit<float>  { static constexpr int exp_terms = 8; };

// fused multiply-add in the arithmetic type itself (no promotion of FP32 to FP64)
__device__ __forceinline__ double fmat(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float  fmat(float a, float b, float c)  { return __builtin_fmaf(a, b, c); }

__device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float  dsqrt(float x)  { return __builtin_sqrtf(x); }
__device__ __forceinline__ double dsin(double x) { return sin(x); }
__device__ __forceinline__ float  dsin(float x)  { return sinf(x); }
__device__ __forceinline__ double dcos(double x) { return cos(x); }
__device__ __forceinline__ float  dcos(float x)  { return cosf(x); }

// ============================================================== SU(2)
template <typename T> struct SU2 {
    static constexpr int NS = 4, NE = 3, NC = 2;
    template <typename U> using rebind = SU2<U>;
    struct Mat { T v[4]; };

    __device__ static __forceinline__ Mat unit() { Mat m; m.v[0] = m.v[1] = m.v[2] = T(0); m.v[3] = T(1); return m; }

    // (a3 + i a.s)(b3 + i b.s) = (a3b3 - a.b) + i(a3 b + b3 a - a x b).s
    __device__ static __forceinline__ Mat mul(const Mat& A, const Mat& B) {
        Mat C;
        C.v[0] = A.v[3]*B.v[0] + B.v[3]*A.v[0] - (A.v[1]*B.v[2] - A.v[2]*B.v[1]);
        C.v[1] = A.v[3]*B.v[1] + B.v[3]*A.v[1] - (A.v[2]*B.v[0] - A.v[0]*B.v[2]);
        C.v[2] = A.v[3]*B.v[2] + B.v[3]*A.v[2] - (A.v[0]*B.v[1] - A.v[1]*B.v[0]);
        C.v[3] = A.v[3]*B.v[3] - A.v[0]*B.v[0] - A.v[1]*B.v[1] - A.v[2]*B.v[2];
        return C;
    }
    __device__ static __forceinline__ Mat dag(const Mat& A) { Mat D; D.v[0] = -A.v[0]; D.v[1] = -A.v[1]; D.v[2] = -A.v[2]; D.v[3] = A.v[3]; return D; }
    __device__ static __forceinline__ Mat mul_nd(const Mat& A, const Mat& B) { return mul(A, dag(B)); }
    __device__ static __forceinline__ Mat mul_dn(const Mat& A, const Mat& B) { return mul(dag(A), B); }
    __device__ static __forceinline__ Mat mul_dd(const Mat& A, const Mat& B) { return mul(dag(A), dag(B)); }
    // op(A) op(B) of two group elements (the SU(3) form saves work this way; here it is mul)
    template <bool dA, bool dB>
    __device__ static __forceinline__ Mat mulsu(const Mat& A, const Mat& B) { return mul(dA ? dag(A) : A, dB ? dag(B) : B); }
    __device__ static __forceinline__ void axpy(T s, const Mat& X, Mat& Y) {
#pragma unroll
        for (int k = 0; k < 4; k++) Y.v[k] += s*X.v[k];
    }
    // o[a] = c ReTr(i t^a M) = -c m_a
    __device__ static __forceinline__ void proj(T c, const Mat& M, T* o) { o[0] = -c*M.v[0]; o[1] = -c*M.v[1]; o[2] = -c*M.v[2]; }
    // o[a] = c ReTr(i t^a A B)
    __device__ static __forceinline__ void proj_mul(T c, const Mat& A, const Mat& B, T* o) {
        o[0] = -c*(A.v[3]*B.v[0] + B.v[3]*A.v[0] - (A.v[1]*B.v[2] - A.v[2]*B.v[1]));
        o[1] = -c*(A.v[3]*B.v[1] + B.v[3]*A.v[1] - (A.v[2]*B.v[0] - A.v[0]*B.v[2]));
        o[2] = -c*(A.v[3]*B.v[2] + B.v[3]*A.v[2] - (A.v[0]*B.v[1] - A.v[1]*B.v[0]));
    }
    // o[a] += c ReTr(i t^a A B): only the vector part of the product
    __device__ static __forceinline__ void proj_mul_acc(T c, const Mat& A, const Mat& B, T* o) {
        o[0] -= c*(A.v[3]*B.v[0] + B.v[3]*A.v[0] - (A.v[1]*B.v[2] - A.v[2]*B.v[1]));
        o[1] -= c*(A.v[3]*B.v[1] + B.v[3]*A.v[1] - (A.v[2]*B.v[0] - A.v[0]*B.v[2]));
        o[2] -= c*(A.v[3]*B.v[2] + B.v[3]*A.v[2] - (A.v[0]*B.v[1] - A.v[1]*B.v[0]));
    }
    // a A + b B
    __device__ static __forceinline__ Mat lin(T a, const Mat& A, T b, const Mat& B) {
        Mat C;
#pragma unroll
        for (int k = 0; k < 4; k++) C.v[k] = a*A.v[k] + b*B.v[k];
        return C;
    }
    // A + b B
    __device__ static __forceinline__ Mat lin1(const Mat& A, T b, const Mat& B) {
        Mat C;
#pragma unroll
        for (int k = 0; k < 4; k++) C.v[k] = fmat(b, B.v[k], A.v[k]);
        return C;
    }
    __device__ static __forceinline__ T retr_1m(const Mat& M) { return T(2)*(T(1) - M.v[3]); }

    // exp(i c al^a sigma_a/2) = cos h + i (c sin(h)/(2h)) al.sigma, h = |c al|/2.
    // FP32: cos h and sin(h)/h as series in s = h^2; for the small angles of
    // the update steps (s <= 0.6) the truncated series is below FP32 rounding
    // and needs no range reduction, square root or division.  FP64 and larger
    // angles: the library's sin and cos (the FP64 form follows the oracle's
    // arithmetic to the last bits that the Gauss-law restoration's stopping
    // test can see).
    __device__ static __forceinline__ Mat expi(T c, const T* al) {
        T x0 = c*al[0], x1 = c*al[1], x2 = c*al[2];
        const T dd = x0*x0 + x1*x1 + x2*x2;
        const T s = T(0.25)*dd;
        constexpr int KC = 5, KS = 4;     // last series terms s^K
        Mat U;
        if (sizeof(T) == 4 && s <= T(0.6)) {
            // Horner: cos h = sum_k (-1)^k s^k/(2k)!, sin(h)/h = sum_k (-1)^k s^k/(2k+1)!
            T ch = T((KC & 1) ? -kInvFact[2*KC] : kInvFact[2*KC]);
#pragma unroll
            for (int k = KC - 1; k >= 0; k--) ch = fmat(ch, s, T((k & 1) ? -kInvFact[2*k] : kInvFact[2*k]));
            T sh = T((KS & 1) ? -kInvFact[2*KS + 1] : kInvFact[2*KS + 1]);
#pragma unroll
            for (int k = KS - 1; k >= 0; k--) sh = fmat(sh, s, T((k & 1) ? -kInvFact[2*k + 1] : kInvFact[2*k + 1]));
            const T m = T(0.5)*c*sh;
            U.v[0] = m*al[0]; U.v[1] = m*al[1]; U.v[2] = m*al[2]; U.v[3] = ch;
        } else if (dd != T(0)) {
            const T d = dsqrt(dd), h = T(0.5)*d, m = c*dsin(h)/d;
            U.v[0] = m*al[0]; U.v[1] = m*al[1]; U.v[2] = m*al[2]; U.v[3] = dcos(h);
        } else {
            U = unit();
        }
        return U;
    }
    // al with exp(i c al t) = U, principal branch (atan2)
    __device__ static __forceinline__ void logi(T c, const Mat& U, T* al) {
        T d = dsqrt(U.v[0]*U.v[0] + U.v[1]*U.v[1] + U.v[2]*U.v[2]);
        if (d != T(0)) {
            T m = T(2)*atan2(d, U.v[3])/(c*d);
            al[0] = m*U.v[0]; al[1] = m*U.v[1]; al[2] = m*U.v[2];
        } else al[0] = al[1] = al[2] = T(0);
    }
    // (V E V^+)^a = (w^2-|v|^2)E + 2v(v.E) + 2w(E x v)
    __device__ static __forceinline__ void adj_act(const Mat& V, const T* E, T* o) {
        T w = V.v[3], vv = V.v[0]*V.v[0] + V.v[1]*V.v[1] + V.v[2]*V.v[2];
        T ve = V.v[0]*E[0] + V.v[1]*E[1] + V.v[2]*E[2], s = w*w - vv;
        T x0 = E[1]*V.v[2] - E[2]*V.v[1], x1 = E[2]*V.v[0] - E[0]*V.v[2], x2 = E[0]*V.v[1] - E[1]*V.v[0];
        o[0] = s*E[0] + T(2)*V.v[0]*ve + T(2)*w*x0;
        o[1] = s*E[1] + T(2)*V.v[1]*ve + T(2)*w*x1;
        o[2] = s*E[2] + T(2)*V.v[2]*ve + T(2)*w*x2;
    }
    __device__ static __forceinline__ void inv_adj_act(const Mat& V, const T* E, T* o) { adj_act(dag(V), E, o); }
    template <bool NEG = false>
    __device__ static __forceinline__ void ad_dag(const Mat& V, const T* f, T* o) {
        adj_act(dag(V), f, o);
        if constexpr (NEG) { o[0] = -o[0]; o[1] = -o[1]; o[2] = -o[2]; }
    }
    __device__ static __forceinline__ T unitarity_norm(const Mat& U) {
        Mat C = mul_nd(U, U);
        return dsqrt(C.v[0]*C.v[0] + C.v[1]*C.v[1] + C.v[2]*C.v[2] + (C.v[3] - T(1))*(C.v[3] - T(1)));
    }
    // normalisation of the quaternion, in Reunit<T> arithmetic
    __device__ static __forceinline__ void reunitarize(Mat& U) {
        using R = Reunit<T>;
        R a = U.v[0], b = U.v[1], c = U.v[2], d = U.v[3];
        R n = rsqrt_reunit<T>(a*a + b*b + c*c + d*d);
        U.v[0] = T(a*n); U.v[1] = T(b*n); U.v[2] = T(c*n); U.v[3] = T(d*n);
    }
    // the SU(3) two-column forms reduce to the whole quaternion
    __device__ static __forceinline__ Mat mul01(const Mat& A, const Mat& B) { return mul(A, B); }
    __device__ static __forceinline__ void reunitarize2(Mat& U) { reunitarize(U); }
    __device__ static __forceinline__ void complete(Mat&) {}
    __device__ static __forceinline__ T fstruct(int a, int b, int c) { return T(((a-b)*(a-c)*(c-b))/2); }
};

// ============================================================== SU(3)
template <typename T> struct SU3 {
    static constexpr int NS = 18, NE = 8, NC = 3;
    template <typename U> using rebind = SU3<U>;
    struct Mat { T r[9], i[9]; };   // element (row, col) at row + 3 col

    __device__ static __forceinline__ Mat unit() {
        Mat m;
#pragma unroll
        for (int k = 0; k < 9; k++) { m.r[k] = T(0); m.i[k] = T(0); }
        m.r[0] = m.r[4] = m.r[8] = T(1);
        return m;
    }
    // C = op(A) op(B); dA/dB select the hermitian conjugate
    template <bool dA, bool dB>
    __device__ static __forceinline__ Mat mulg(const Mat& A, const Mat& B) {
        Mat C;
#pragma unroll
        for (int row = 0; row < 3; row++) {
#pragma unroll
            for (int col = 0; col < 3; col++) {
                T sr = T(0), si = T(0);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const int ia = dA ? (k + 3*row) : (row + 3*k);
                    const int ib = dB ? (col + 3*k) : (k + 3*col);
                    const T ar = A.r[ia], ai = dA ? -A.i[ia] : A.i[ia];
                    const T br = B.r[ib], bi = dB ? -B.i[ib] : B.i[ib];
                    sr = fmat(ar, br, sr); sr = fmat(-ai, bi, sr);
                    si = fmat(ar, bi, si); si = fmat(ai, br, si);
                }
                C.r[row + 3*col] = sr; C.i[row + 3*col] = si;
            }
        }
        return C;
    }
    __device__ static __forceinline__ Mat mul(const Mat& A, const Mat& B) { return mulg<false, false>(A, B); }
    __device__ static __forceinline__ Mat mul_nd(const Mat& A, const Mat& B) { return mulg<false, true>(A, B); }
    __device__ static __forceinline__ Mat mul_dn(const Mat& A, const Mat& B) { return mulg<true, false>(A, B); }
    __device__ static __forceinline__ Mat mul_dd(const Mat& A, const Mat& B) { return mulg<true, true>(A, B); }
    // column 2 = conj(col0 x col1) from columns 0 and 1 (exact for special unitary m)
    __device__ static __forceinline__ void complete(Mat& m) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int j = (i + 1) % 3, l = (i + 2) % 3;
            // c0[j] c1[l] - c0[l] c1[j], conjugated
            const T ar = m.r[j], ai = m.i[j], br = m.r[3 + l], bi = m.i[3 + l];
            const T cr = m.r[l], ci = m.i[l], dr = m.r[3 + j], di = m.i[3 + j];
            m.r[6 + i] = fmat(ar, br, fmat(-ai, bi, fmat(-cr, dr, ci*di)));
            m.i[6 + i] = -fmat(ar, bi, fmat(ai, br, fmat(-cr, di, -(ci*dr))));
        }
    }
    // op(A) op(B) of two special unitary matrices: columns 0 and 1 by the
    // product (72 multiply-adds), column 2 rebuilt (24) instead of multiplied
    // (36).  op(B)'s column 2 is not read, so for dB = false a B from ld2
    // needs no rebuild at all, and for dB = true only two of its three
    // rebuilt entries are read; the compiler drops the unread ones.
    template <bool dA, bool dB>
    __device__ static __forceinline__ Mat mulsu(const Mat& A, const Mat& B) {
        Mat C;
#pragma unroll
        for (int row = 0; row < 3; row++) {
#pragma unroll
            for (int col = 0; col < 2; col++) {
                T sr = T(0), si = T(0);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const int ia = dA ? (k + 3*row) : (row + 3*k);
                    const int ib = dB ? (col + 3*k) : (k + 3*col);
                    const T ar = A.r[ia], ai = dA ? -A.i[ia] : A.i[ia];
                    const T br = B.r[ib], bi = dB ? -B.i[ib] : B.i[ib];
                    sr = fmat(ar, br, sr); sr = fmat(-ai, bi, sr);
                    si = fmat(ar, bi, si); si = fmat(ai, br, si);
                }
                C.r[row + 3*col] = sr; C.i[row + 3*col] = si;
            }
        }
        complete(C);
        return C;
    }
    __device__ static __forceinline__ Mat dag(const Mat& A) {
        Mat D;
#pragma unroll
        for (int row = 0; row < 3; row++)
#pragma unroll
            for (int col = 0; col < 3; col++) { D.r[row + 3*col] = A.r[col + 3*row]; D.i[row + 3*col] = -A.i[col + 3*row]; }
        return D;
    }
    __device__ static __forceinline__ void axpy(T s, const Mat& X, Mat& Y) {
#pragma unroll
        for (int k = 0; k < 9; k++) { Y.r[k] += s*X.r[k]; Y.i[k] += s*X.i[k]; }
    }
    // o[a] = c ReTr(i t^a M), Gell-Mann order
    __device__ static __forceinline__ void proj(T c, const Mat& M, T* o) {
        const T h = T(-0.5)*c, s3 = T(0.57735026918962576450914878050196);
        o[0] = h*(M.i[3] + M.i[1]);
        o[1] = h*(M.r[3] - M.r[1]);
        o[2] = h*(M.i[0] - M.i[4]);
        o[3] = h*(M.i[6] + M.i[2]);
        o[4] = h*(M.r[6] - M.r[2]);
        o[5] = h*(M.i[7] + M.i[5]);
        o[6] = h*(M.r[7] - M.r[5]);
        o[7] = h*(M.i[0] + M.i[4] - T(2)*M.i[8])*s3;
    }
    __device__ static __forceinline__ T retr_1m(const Mat& M) { return T(3) - M.r[0] - M.r[4] - M.r[8]; }
    // o[a] += c ReTr(i t^a A B), evaluating only the entries of A B the
    // projection reads: the six off-diagonal ones and the diagonal's imaginary parts
    // (ACC false: o[a] = c ReTr(i t^a A B))
    template <bool ACC = true>
    __device__ static __forceinline__ void proj_mul_acc(T c, const Mat& A, const Mat& B, T* o) {
        auto dim = [&](int d) -> T {
            T si = T(0);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                si = fmat(A.r[d + 3*k], B.i[k + 3*d], si);
                si = fmat(A.i[d + 3*k], B.r[k + 3*d], si);
            }
            return si;
        };
        // for an off-diagonal pair (r, c) the projection reads only
        // Re X_rc - Re X_cr and Im X_rc + Im X_cr (X = A B): one multiply-add
        // chain each, over the twelve products of both entries
        auto pair = [&](int r, int c, T& dre, T& sim) {
            T a = T(0), b = T(0);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int ir = r + 3*k, jc = k + 3*c, ic = c + 3*k, jr = k + 3*r;
                a = fmat(A.r[ir], B.r[jc], a); a = fmat(-A.i[ir], B.i[jc], a);
                a = fmat(-A.r[ic], B.r[jr], a); a = fmat(A.i[ic], B.i[jr], a);
                b = fmat(A.r[ir], B.i[jc], b); b = fmat(A.i[ir], B.r[jc], b);
                b = fmat(A.r[ic], B.i[jr], b); b = fmat(A.i[ic], B.r[jr], b);
            }
            dre = a; sim = b;
        };
        T d01, s01, d02, s02, d12, s12;
        pair(0, 1, d01, s01); pair(0, 2, d02, s02); pair(1, 2, d12, s12);
        const T i00 = dim(0), i11 = dim(1), i22 = dim(2);
        const T h = T(-0.5)*c, s3 = T(0.57735026918962576450914878050196);
        const T p[8] = {h*s01, h*d01, h*(i00 - i11), h*s02,
                        h*d02, h*s12, h*d12, h*(i00 + i11 - T(2)*i22)*s3};
#pragma unroll
        for (int a = 0; a < 8; a++) o[a] = ACC ? o[a] + p[a] : p[a];
    }
    __device__ static __forceinline__ void proj_mul(T c, const Mat& A, const Mat& B, T* o) { proj_mul_acc<false>(c, A, B, o); }
    // a A + b B
    __device__ static __forceinline__ Mat lin(T a, const Mat& A, T b, const Mat& B) {
        Mat C;
#pragma unroll
        for (int k = 0; k < 9; k++) { C.r[k] = a*A.r[k] + b*B.r[k]; C.i[k] = a*A.i[k] + b*B.i[k]; }
        return C;
    }
    // A + b B
    __device__ static __forceinline__ Mat lin1(const Mat& A, T b, const Mat& B) {
        Mat C;
#pragma unroll
        for (int k = 0; k < 9; k++) { C.r[k] = fmat(b, B.r[k], A.r[k]); C.i[k] = fmat(b, B.i[k], A.i[k]); }
        return C;
    }

    // hermitian H = al^a t^a
    __device__ static __forceinline__ Mat alg_matrix(T c, const T* al) {
        const T s3 = T(0.57735026918962576450914878050196), h = T(0.5)*c;
        Mat A;
        A.r[0] = h*(al[2] + al[7]*s3);  A.i[0] = T(0);
        A.r[4] = -h*(al[2] - al[7]*s3); A.i[4] = T(0);
        A.r[8] = -c*al[7]*s3;           A.i[8] = T(0);
        A.r[1] = h*al[0]; A.i[1] = h*al[1];     // (1,0)
        A.r[2] = h*al[3]; A.i[2] = h*al[4];     // (2,0)
        A.r[5] = h*al[5]; A.i[5] = h*al[6];     // (2,1)
        A.r[3] = A.r[1]; A.i[3] = -A.i[1];      // (0,1)
        A.r[6] = A.r[2]; A.i[6] = -A.i[2];      // (0,2)
        A.r[7] = A.r[5]; A.i[7] = -A.i[5];      // (1,2)
        return A;
    }
    // Q^2 of a hermitian Q with a real diagonal d and lower triangle
    // q10, q20, q21 (upper triangle by conjugation): the diagonal and the
    // lower triangle of Q^2, no product with a vanishing imaginary part
    //   (Q^2)_10 = q10 (d0 + d1) + conj(q21) q20
    //   (Q^2)_20 = q20 (d0 + d2) + q21 q10
    //   (Q^2)_21 = q21 (d1 + d2) + q20 conj(q10)
    __device__ static __forceinline__ Mat herm_sq(const Mat& Q) {
        Mat P;
        const T d0 = Q.r[0], d1 = Q.r[4], d2 = Q.r[8];
        const T ar = Q.r[1], ai = Q.i[1], br = Q.r[2], bi = Q.i[2], cr = Q.r[5], ci = Q.i[5];   // q10, q20, q21
        const T na = fmat(ar, ar, ai*ai), nb = fmat(br, br, bi*bi), nc = fmat(cr, cr, ci*ci);
        P.r[0] = fmat(d0, d0, na + nb);
        P.r[4] = fmat(d1, d1, na + nc);
        P.r[8] = fmat(d2, d2, nb + nc);
        P.i[0] = T(0); P.i[4] = T(0); P.i[8] = T(0);
        const T s01 = d0 + d1, s02 = d0 + d2, s12 = d1 + d2;
        P.r[1] = fmat(ar, s01, fmat(cr, br, ci*bi));
        P.i[1] = fmat(ai, s01, fmat(cr, bi, -(ci*br)));
        P.r[2] = fmat(br, s02, fmat(cr, ar, -(ci*ai)));
        P.i[2] = fmat(bi, s02, fmat(cr, ai, ci*ar));
        P.r[5] = fmat(cr, s12, fmat(br, ar, bi*ai));
        P.i[5] = fmat(ci, s12, fmat(bi, ar, -(br*ai)));
        P.r[3] = P.r[1]; P.i[3] = -P.i[1];
        P.r[6] = P.r[2]; P.i[6] = -P.i[2];
        P.r[7] = P.r[5]; P.i[7] = -P.i[5];
        return P;
    }
    // o^a = tr(lambda_a X) for hermitian X
    __device__ static __forceinline__ void herm_alg(const Mat& X, T* o) {
        const T s3 = T(0.57735026918962576450914878050196);
        o[0] = X.r[3] + X.r[1];
        o[1] = X.i[1] - X.i[3];
        o[2] = X.r[0] - X.r[4];
        o[3] = X.r[6] + X.r[2];
        o[4] = X.i[2] - X.i[6];
        o[5] = X.r[7] + X.r[5];
        o[6] = X.i[5] - X.i[7];
        o[7] = (X.r[0] + X.r[4] - T(2)*X.r[8])*s3;
    }
    __device__ static __forceinline__ void adj_act(const Mat& V, const T* E, T* o) {
        Mat A = alg_matrix(T(1), E);
        Mat X = mul_nd(mul(V, A), V);
        herm_alg(X, o);
    }
    __device__ static __forceinline__ void inv_adj_act(const Mat& V, const T* E, T* o) {
        Mat A = alg_matrix(T(1), E);
        Mat X = mul(mul_dn(V, A), V);
        herm_alg(X, o);
    }
    // o = (V^+ H V)^a, H = f^a t^a, for special unitary V.  With H' = 2H (the
    // Gell-Mann combination f^a lambda_a, entries read off f) and
    // X' = V^+ H' V: o_1..6 are the real and imaginary parts of the lower
    // triangle of X', o_3 and o_8 come from its diagonal, and X'(2,2) =
    // -(X'(0,0) + X'(1,1)) since H' is traceless, so only columns 0 and 1
    // of Y = H' V are formed (60 instead of 90 multiply-adds)
    // NEG: o = -(V^+ H V), the sign taken into H' (free operand negations)
    template <bool NEG = false>
    __device__ static __forceinline__ void ad_dag(const Mat& V, const T* f0, T* o) {
        const T s3 = T(0.57735026918962576450914878050196);
        T f[8];
#pragma unroll
        for (int a = 0; a < 8; a++) f[a] = NEG ? -f0[a] : f0[a];
        // H' = f^a lambda_a: real diagonal, lower triangle (1,0), (2,0), (2,1)
        const T h00 = fmat(f[7], s3, f[2]), h11 = fmat(f[7], s3, -f[2]), h22 = T(-2)*f[7]*s3;
        const T hr[3][3] = {{h00, f[0], f[3]}, {f[0], h11, f[5]}, {f[3], f[5], h22}};
        const T hi[3][3] = {{T(0), -f[1], -f[4]}, {f[1], T(0), -f[6]}, {f[4], f[6], T(0)}};
        T Yr[3][2], Yi[3][2];
#pragma unroll
        for (int r = 0; r < 3; r++) {
#pragma unroll
            for (int c = 0; c < 2; c++) {
                T sr = hr[r][r]*V.r[r + 3*c], si = hr[r][r]*V.i[r + 3*c];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    if (k == r) continue;
                    sr = fmat(hr[r][k], V.r[k + 3*c], sr); sr = fmat(-hi[r][k], V.i[k + 3*c], sr);
                    si = fmat(hr[r][k], V.i[k + 3*c], si); si = fmat(hi[r][k], V.r[k + 3*c], si);
                }
                Yr[r][c] = sr; Yi[r][c] = si;
            }
        }
        // X'(r, c) = sum_k conj(V(k, r)) Y(k, c)
        auto xre = [&](int r, int c) {
            T s = T(0);
#pragma unroll
            for (int k = 0; k < 3; k++) { s = fmat(V.r[k + 3*r], Yr[k][c], s); s = fmat(V.i[k + 3*r], Yi[k][c], s); }
            return s;
        };
        auto xim = [&](int r, int c) {
            T s = T(0);
#pragma unroll
            for (int k = 0; k < 3; k++) { s = fmat(V.r[k + 3*r], Yi[k][c], s); s = fmat(-V.i[k + 3*r], Yr[k][c], s); }
            return s;
        };
        const T d0 = xre(0, 0), d1 = xre(1, 1);
        o[0] = xre(1, 0);
        o[1] = xim(1, 0);
        o[2] = T(0.5)*(d0 - d1);
        o[3] = xre(2, 0);
        o[4] = xim(2, 0);
        o[5] = xre(2, 1);
        o[6] = xim(2, 1);
        o[7] = T(0.86602540378443864676372317075294)*(d0 + d1);
    }

    // exp(i Q), Q = c al^a t^a, as f0 + f1 Q + f2 Q^2 (Cayley-Hamilton):
    // Q^3 = c1 Q + c0 with c1 = tr(Q^2)/2, c0 = det Q.  The Taylor series of
    // exp(iQ) is summed in the (1, Q, Q^2) basis with scalar recurrences, after
    // scaling Q by 2^-s so that c1 <= 1/4; the result is squared s times in
    // the same basis.
    __device__ static __forceinline__ Mat expi(T c, const T* al) {
        const Mat Q = alg_matrix(c, al);
        // Q^2 first: c1 = tr(Q^2)/2 and, Q being traceless, c0 = det Q =
        // tr(Q^3)/3 = (sum_i Q_ii (Q^2)_ii + 2 sum_{i>j} Re(Q_ij conj((Q^2)_ij)))/3
        const Mat Q2 = herm_sq(Q);
        const T c1 = T(0.5)*(Q2.r[0] + Q2.r[4] + Q2.r[8]);
        if (c1 == T(0)) return unit();
        T t3 = fmat(Q.r[1], Q2.r[1], Q.i[1]*Q2.i[1]);
        t3 = fmat(Q.r[2], Q2.r[2], t3); t3 = fmat(Q.i[2], Q2.i[2], t3);
        t3 = fmat(Q.r[5], Q2.r[5], t3); t3 = fmat(Q.i[5], Q2.i[5], t3);
        const T d3 = fmat(Q.r[0], Q2.r[0], fmat(Q.r[4], Q2.r[4], Q.r[8]*Q2.r[8]));
        const T c0 = fmat(T(2), t3, d3)*T(1.0/3.0);
        int s = 0;
        T cs1 = c1, cs0 = c0, sc = T(1);
        while (cs1 > T(0.25) && s < 40) { cs1 *= T(0.25); cs0 *= T(0.125); sc *= T(0.5); s++; }
        // Taylor series in the (1,Q,Q^2) basis.  Terms 0 .. 4 written out
        // (Q^0 .. Q^4 = 1, Q, Q^2, cs0 + cs1 Q, cs0 Q + cs1 Q^2), so that no
        // coefficient is multiplied by a vanishing basis component
        T f0r = T(1), f0i = T(-kInvFact[3])*cs0, f1r = T(kInvFact[4])*cs0;
        T f1i = fmat(T(-kInvFact[3]), cs1, T(1)), f2r = fmat(T(kInvFact[4]), cs1, T(-kInvFact[2])), f2i = T(0);
        T pa = cs0*cs1, pb = cs1*cs1, pd = cs0;         // Q^5 = pa + pb Q + pd Q^2
#pragma unroll
        for (int k = 5; k <= Lit<T>::exp_terms; k++) {
            const T fact = T(kInvFact[k]);
            const T w = ((k >> 1) & 1) ? -fact : fact;   // i^k = (+1, +i, -1, -i)
            if (k & 1) { f0i = fmat(w, pa, f0i); f1i = fmat(w, pb, f1i); f2i = fmat(w, pd, f2i); }
            else       { f0r = fmat(w, pa, f0r); f1r = fmat(w, pb, f1r); f2r = fmat(w, pd, f2r); }
            if (k < Lit<T>::exp_terms) {
                const T na = cs0*pd, nb = fmat(cs1, pd, pa), nd = pb;
                pa = na; pb = nb; pd = nd;
            }
        }
        for (int k = 0; k < s; k++) {
            // p^2 reduced with Q^3 = cs1 Q + cs0, Q^4 = cs1 Q^2 + cs0 Q
            T a0r = f0r*f0r - f0i*f0i, a0i = T(2)*f0r*f0i;
            T m12r = f1r*f2r - f1i*f2i, m12i = f1r*f2i + f1i*f2r;
            T m22r = f2r*f2r - f2i*f2i, m22i = T(2)*f2r*f2i;
            T m01r = f0r*f1r - f0i*f1i, m01i = f0r*f1i + f0i*f1r;
            T m02r = f0r*f2r - f0i*f2i, m02i = f0r*f2i + f0i*f2r;
            T m11r = f1r*f1r - f1i*f1i, m11i = T(2)*f1r*f1i;
            T n0r = a0r + T(2)*m12r*cs0, n0i = a0i + T(2)*m12i*cs0;
            T n1r = T(2)*m01r + T(2)*m12r*cs1 + m22r*cs0, n1i = T(2)*m01i + T(2)*m12i*cs1 + m22i*cs0;
            T n2r = T(2)*m02r + m11r + m22r*cs1, n2i = T(2)*m02i + m11i + m22i*cs1;
            f0r = n0r; f0i = n0i; f1r = n1r; f1i = n1i; f2r = n2r; f2i = n2i;
        }
        // U = f0 + f1 Qs + f2 Qs^2 with Qs = sc Q, as f0 + (sc f1) Q + (sc^2 f2) Q^2
        f1r *= sc; f1i *= sc;
        f2r *= sc*sc; f2i *= sc*sc;
        Mat U;
        // diagonal: Q and Q^2 real there
#pragma unroll
        for (int k = 0; k < 9; k += 4) {
            U.r[k] = fmat(f1r, Q.r[k], fmat(f2r, Q2.r[k], f0r));
            U.i[k] = fmat(f1i, Q.r[k], fmat(f2i, Q2.r[k], f0i));
        }
        // off-diagonal pairs (lower l, upper u): Q_u = conj(Q_l), Q2_u = conj(Q2_l),
        // so U_l = (A - B) + i (C + D) and U_u = (A + B) + i (C - D) with
        // A = f1r q.r + f2r p.r, B = f1i q.i + f2i p.i, C = f1i q.r + f2i p.r, D = f1r q.i + f2r p.i
        constexpr int lo[3] = {1, 2, 5}, up[3] = {3, 6, 7};
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int l = lo[j], u = up[j];
            const T a = fmat(f1r, Q.r[l], f2r*Q2.r[l]), b = fmat(f1i, Q.i[l], f2i*Q2.i[l]);
            const T c = fmat(f1i, Q.r[l], f2i*Q2.r[l]), d = fmat(f1r, Q.i[l], f2r*Q2.i[l]);
            U.r[l] = a - b; U.r[u] = a + b;
            U.i[l] = c + d; U.i[u] = c - d;
        }
        return U;
    }

    __device__ static __forceinline__ T unitarity_norm(const Mat& U) {
        Mat C = mul_nd(U, U);
        T n = T(0);
#pragma unroll
        for (int k = 0; k < 9; k++) {
            T d = C.r[k] - ((k == 0 || k == 4 || k == 8) ? T(1) : T(0));
            n += d*d + C.i[k]*C.i[k];
        }
        return dsqrt(n);
    }
    // columns 0 and 1 of op(A) B (column 2 of the result is left zero): the
    // part of a link update that the reunitarisation and a two-column store read
    __device__ static __forceinline__ Mat mul01(const Mat& A, const Mat& B) {
        Mat C;
#pragma unroll
        for (int row = 0; row < 3; row++) {
#pragma unroll
            for (int col = 0; col < 2; col++) {
                T sr = T(0), si = T(0);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const int ia = row + 3*k, ib = k + 3*col;
                    sr = fmat(A.r[ia], B.r[ib], sr); sr = fmat(-A.i[ia], B.i[ib], sr);
                    si = fmat(A.r[ia], B.i[ib], si); si = fmat(A.i[ia], B.r[ib], si);
                }
                C.r[row + 3*col] = sr; C.i[row + 3*col] = si;
            }
            C.r[row + 6] = T(0); C.i[row + 6] = T(0);
        }
        return C;
    }
    // Gram-Schmidt on columns 0 and 1 in Reunit<T> arithmetic; with full,
    // column 2 = conj(col0 x col1) in the same arithmetic, else untouched
    template <bool full>
    __device__ static __forceinline__ void reunitarize_cols(Mat& U) {
        using R = Reunit<T>;
        R c0r[3], c0i[3], c1r[3], c1i[3];
#pragma unroll
        for (int j = 0; j < 3; j++) { c0r[j] = U.r[j]; c0i[j] = U.i[j]; c1r[j] = U.r[3 + j]; c1i[j] = U.i[3 + j]; }
        // sums as single multiply-add chains
        R n = c0r[0]*c0r[0];
        n = fmat(c0i[0], c0i[0], n);
#pragma unroll
        for (int j = 1; j < 3; j++) { n = fmat(c0r[j], c0r[j], n); n = fmat(c0i[j], c0i[j], n); }
        n = rsqrt_reunit<T>(n);
#pragma unroll
        for (int j = 0; j < 3; j++) { c0r[j] *= n; c0i[j] *= n; }
        R pr = c0r[0]*c1r[0], pi = c0r[0]*c1i[0];   // <c0, c1>
        pr = fmat(c0i[0], c1i[0], pr); pi = fmat(-c0i[0], c1r[0], pi);
#pragma unroll
        for (int j = 1; j < 3; j++) {
            pr = fmat(c0r[j], c1r[j], pr); pr = fmat(c0i[j], c1i[j], pr);
            pi = fmat(c0r[j], c1i[j], pi); pi = fmat(-c0i[j], c1r[j], pi);
        }
#pragma unroll
        for (int j = 0; j < 3; j++) {
            c1r[j] = fmat(-pr, c0r[j], fmat(pi, c0i[j], c1r[j]));
            c1i[j] = fmat(-pr, c0i[j], fmat(-pi, c0r[j], c1i[j]));
        }
        n = c1r[0]*c1r[0];
        n = fmat(c1i[0], c1i[0], n);
#pragma unroll
        for (int j = 1; j < 3; j++) { n = fmat(c1r[j], c1r[j], n); n = fmat(c1i[j], c1i[j], n); }
        n = rsqrt_reunit<T>(n);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            c1r[j] *= n; c1i[j] *= n;
            U.r[j] = T(c0r[j]); U.i[j] = T(c0i[j]);
            U.r[3 + j] = T(c1r[j]); U.i[3 + j] = T(c1i[j]);
        }
        if constexpr (full) {
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const int a = (i + 1) % 3, b = (i + 2) % 3;
                const R cr = c0r[a]*c1r[b] - c0i[a]*c1i[b] - (c0r[b]*c1r[a] - c0i[b]*c1i[a]);
                const R ci = c0r[a]*c1i[b] + c0i[a]*c1r[b] - (c0r[b]*c1i[a] + c0i[b]*c1r[a]);
                U.r[6 + i] = T(cr); U.i[6 + i] = T(-ci);
            }
        }
    }
    // the nearest-by-Gram-Schmidt special unitary matrix: columns 0, 1
    // orthonormalised, column 2 = conj(col0 x col1)
    __device__ static __forceinline__ void reunitarize(Mat& U) { reunitarize_cols<true>(U); }
    // columns 0 and 1 only (a link stored as two columns, or completed later)
    __device__ static __forceinline__ void reunitarize2(Mat& U) { reunitarize_cols<false>(U); }
    __device__ static T fstruct(int a, int b, int c);
};

// ---------------------------------------------------------------- SU(3) log (FP64)
// -i log W for W in SU(3): eigenvectors from a complex Jacobi diagonalisation
// of the hermitian Cayley transform C = i(1 - W')(1 + W')^-1 of W' = e^{-i s} W
// (the shift s keeps every eigenvalue of W' away from -1), eigen-phases
// phi = 2 atan(c) + s wrapped to the principal branch (-pi, pi], then
// projected onto the generators.
struct Log3 {
    struct C { double re, im; };
    __device__ static __forceinline__ C cm(C a, C b) { return {a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re}; }
    __device__ static __forceinline__ C cs(C a, C b) { return {a.re - b.re, a.im - b.im}; }
    __device__ static __forceinline__ C ca(C a, C b) { return {a.re + b.re, a.im + b.im}; }
    __device__ static __forceinline__ C cj(C a) { return {a.re, -a.im}; }
    __device__ static __forceinline__ C csc(double s, C a) { return {s*a.re, s*a.im}; }

    __device__ static __forceinline__ void mm(const C* A, const C* B, C* R) {
        C T[9];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                C s = {0, 0};
#pragma unroll
                for (int k = 0; k < 3; k++) s = ca(s, cm(A[r + 3*k], B[k + 3*c]));
                T[r + 3*c] = s;
            }
#pragma unroll
        for (int k = 0; k < 9; k++) R[k] = T[k];
    }
    __device__ static __forceinline__ C det3(const C* M) {
        C c00 = cs(cm(M[4], M[8]), cm(M[7], M[5]));
        C c01 = cs(cm(M[7], M[2]), cm(M[1], M[8]));
        C c02 = cs(cm(M[1], M[5]), cm(M[4], M[2]));
        return ca(ca(cm(M[0], c00), cm(M[3], c01)), cm(M[6], c02));
    }
    __device__ static __forceinline__ void inv3(const C* M, C* I) {
        C c00 = cs(cm(M[4], M[8]), cm(M[7], M[5]));
        C c01 = cs(cm(M[7], M[2]), cm(M[1], M[8]));
        C c02 = cs(cm(M[1], M[5]), cm(M[4], M[2]));
        C det = ca(ca(cm(M[0], c00), cm(M[3], c01)), cm(M[6], c02));
        double dd = det.re*det.re + det.im*det.im;
        C id = csc(1.0/dd, cj(det));
        C c10 = cs(cm(M[6], M[5]), cm(M[3], M[8]));
        C c11 = cs(cm(M[0], M[8]), cm(M[6], M[2]));
        C c12 = cs(cm(M[3], M[2]), cm(M[0], M[5]));
        C c20 = cs(cm(M[3], M[7]), cm(M[6], M[4]));
        C c21 = cs(cm(M[6], M[1]), cm(M[0], M[7]));
        C c22 = cs(cm(M[0], M[4]), cm(M[3], M[1]));
        I[0] = cm(c00, id); I[3] = cm(c10, id); I[6] = cm(c20, id);
        I[1] = cm(c01, id); I[4] = cm(c11, id); I[7] = cm(c21, id);
        I[2] = cm(c02, id); I[5] = cm(c12, id); I[8] = cm(c22, id);
    }
    // one Jacobi rotation in the (p, q) plane: A <- R^+ A R, V <- V R, with R
    // the identity but for R_pp = c, R_pq = s, R_qp = -s conj(e),
    // R_qq = c conj(e), e = A_pq/|A_pq|; only columns p, q of A R and V R and
    // rows p, q of R^+ (A R) change
    // 1/sqrt(x) and 1/x for finite x > 0 (x >= 1 for rcp): the hardware
    // estimates refined by Newton steps to within a few ulp (a rotation needs
    // c^2 + s^2 = 1 to rounding, not correctly rounded angles: Jacobi sweeps
    // converge from any rotation close to the exact one)
    __device__ static __forceinline__ double rsq(double x) {
        const double h = 0.5*x;
        double r = __builtin_amdgcn_rsq(x);
        r = r*__builtin_fma(-h*r, r, 1.5);
        return r*__builtin_fma(-h*r, r, 1.5);
    }
    __device__ static __forceinline__ double rcp(double x) {
        double r = __builtin_amdgcn_rcp(x);
        r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
        return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    }
    template <int p, int q>
    __device__ static __forceinline__ void rot(C* A, C* V) {
        const C apq = A[p + 3*q];
        const double g2 = apq.re*apq.re + apq.im*apq.im;
        if (g2 == 0.0) return;
        const double rg = rsq(g2);     // 1/|A_pq|
        const C e = csc(rg, apq), ec = cj(e);
        const double th = (A[q + 3*q].re - A[p + 3*p].re)*(0.5*rg);
        // t = sgn(th)/(|th| + sqrt(th^2 + 1)); for |th| > 2^500 (th^2 would
        // overflow) sqrt(th^2 + 1) = |th| to rounding
        const double ath = __builtin_fabs(th), q2 = th*th + 1.0;
        const double den = ath > 0x1p500 ? ath + ath : ath + q2*rsq(q2);
        const double t = (th >= 0 ? 1.0 : -1.0)*rcp(den);
        const double cc = rsq(t*t + 1.0), sn = t*cc;
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const C u = cm(A[r + 3*q], ec), a = A[r + 3*p];
            A[r + 3*p] = cs(csc(cc, a), csc(sn, u));
            A[r + 3*q] = ca(csc(sn, a), csc(cc, u));
            const C v = cm(V[r + 3*q], ec), b = V[r + 3*p];
            V[r + 3*p] = cs(csc(cc, b), csc(sn, v));
            V[r + 3*q] = ca(csc(sn, b), csc(cc, v));
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const C u = cm(e, A[q + 3*c]), a = A[p + 3*c];
            A[p + 3*c] = cs(csc(cc, a), csc(sn, u));
            A[q + 3*c] = ca(csc(sn, a), csc(cc, u));
        }
        A[p + 3*q] = A[q + 3*p] = {0, 0};
#pragma unroll
        for (int k = 0; k < 3; k++) A[k + 3*k].im = 0.0;
    }
    // hermitian Jacobi; A overwritten, V eigenvectors (columns), w eigenvalues
    __device__ static __forceinline__ void jacobi(C* A, C* V, double* w) {
#pragma unroll
        for (int k = 0; k < 9; k++) V[k] = {0, 0};
        V[0] = V[4] = V[8] = {1, 0};
        for (int sweep = 0; sweep < 12; sweep++) {
            double off = A[3].re*A[3].re + A[3].im*A[3].im + A[6].re*A[6].re + A[6].im*A[6].im + A[7].re*A[7].re + A[7].im*A[7].im;
            double dia = A[0].re*A[0].re + A[4].re*A[4].re + A[8].re*A[8].re;
            if (off == 0.0 || off < 1e-36*dia) break;
            rot<0, 1>(A, V);
            rot<0, 2>(A, V);
            rot<1, 2>(A, V);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) w[k] = A[k + 3*k].re;
    }
    // W given as re/im arrays (row + 3 col); out al with exp(i c al t) = W
    __device__ static void logi(double c, const double* Wr, const double* Wi, double* al) {
        C W[9];
        for (int k = 0; k < 9; k++) W[k] = {Wr[k], Wi[k]};
        // shifts 0, pi/2, -pi/2, pi and their phases e^{-i shift}
        const double shifts[4] = {0.0, 1.5707963267948966, -1.5707963267948966, 3.141592653589793};
        const C phases[4] = {{1.0, 0.0}, {0.0, -1.0}, {0.0, 1.0}, {-1.0, 0.0}};
        double best = -1.0, sft = 0.0;
        C ph = phases[0];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            C M[9];
#pragma unroll
            for (int j = 0; j < 9; j++) M[j] = cm(phases[k], W[j]);
            M[0].re += 1; M[4].re += 1; M[8].re += 1;
            C d = det3(M);
            double ad = d.re*d.re + d.im*d.im;
            if (ad > best) { best = ad; sft = shifts[k]; ph = phases[k]; }
            if (k == 0 && ad > 0.25) break;
        }
        C P[9], Mi[9];
        for (int j = 0; j < 9; j++) { C wp = cm(ph, W[j]); P[j] = wp; Mi[j] = csc(-1.0, wp); }
        for (int j = 0; j < 3; j++) { P[j + 3*j].re += 1; Mi[j + 3*j].re += 1; }
        C Pinv[9], Cm[9], Ch[9];
        inv3(P, Pinv);
        mm(Mi, Pinv, Cm);
        for (int r = 0; r < 3; r++) for (int cc = 0; cc < 3; cc++) {
            C a = cm({0, 1}, Cm[r + 3*cc]), b = cj(cm({0, 1}, Cm[cc + 3*r]));
            Ch[r + 3*cc] = csc(0.5, ca(a, b));
        }
        C V[9]; double w[3];
        jacobi(Ch, V, w);
        double Hr[9], Hi[9];
        for (int k = 0; k < 9; k++) { Hr[k] = 0; Hi[k] = 0; }
        for (int j = 0; j < 3; j++) {
            double phi = 2.0*atan(w[j]) + sft;
            while (phi > 3.141592653589793) phi -= 6.283185307179586;
            while (phi <= -3.141592653589793) phi += 6.283185307179586;
            for (int r = 0; r < 3; r++) for (int cc = 0; cc < 3; cc++) {
                C v = cm(V[r + 3*j], cj(V[cc + 3*j]));
                Hr[r + 3*cc] += phi*v.re; Hi[r + 3*cc] += phi*v.im;
            }
        }
        const double s3 = 0.57735026918962576450914878050196;
        al[0] = (Hr[3] + Hr[1])/c;
        al[1] = (Hi[1] - Hi[3])/c;
        al[2] = (Hr[0] - Hr[4])/c;
        al[3] = (Hr[6] + Hr[2])/c;
        al[4] = (Hi[2] - Hi[6])/c;
        al[5] = (Hr[7] + Hr[5])/c;
        al[6] = (Hi[5] - Hi[7])/c;
        al[7] = (Hr[0] + Hr[4] - 2.0*Hr[8])*s3/c;
    }
};

// su(3) structure constants, 0-based Gell-Mann indices
template <typename T>
__device__ T SU3<T>::fstruct(int a, int b, int c) {
    // sort (a,b,c) keeping track of the permutation sign
    int s = 1;
    if (a > b) { int t = a; a = b; b = t; s = -s; }
    if (b > c) { int t = b; b = c; c = t; s = -s; }
    if (a > b) { int t = a; a = b; b = t; s = -s; }
    T v = T(0);
    const int key = a*64 + b*8 + c;
    switch (key) {
        case 0*64 + 1*8 + 2: v = T(1); break;
        case 0*64 + 3*8 + 6: v = T(0.5); break;
        case 0*64 + 4*8 + 5: v = T(-0.5); break;
        case 1*64 + 3*8 + 5: v = T(0.5); break;
        case 1*64 + 4*8 + 6: v = T(0.5); break;
        case 2*64 + 3*8 + 4: v = T(0.5); break;
        case 2*64 + 5*8 + 6: v = T(-0.5); break;
        case 3*64 + 4*8 + 7: v = T(0.86602540378443864676372317075294); break;
        case 5*64 + 6*8 + 7: v = T(0.86602540378443864676372317075294); break;
        default: v = T(0);
    }
    return s*v;
}

// ---------------------------------------------------------------- Philox-4x32-10
struct Philox {
    __device__ static __forceinline__ void round4(uint32_t* c, const uint32_t* k) {
        const uint64_t p0 = (uint64_t)0xD2511F53u*c[0], p1 = (uint64_t)0xCD9E8D57u*c[2];
        const uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0, h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
        const uint32_t n0 = h1 ^ c[1] ^ k[0], n2 = h0 ^ c[3] ^ k[1];
        c[0] = n0; c[1] = l1; c[2] = n2; c[3] = l0;
    }
    __device__ static __forceinline__ void gen(uint32_t* c, uint32_t k0, uint32_t k1) {
        uint32_t k[2] = {k0, k1};
#pragma unroll
        for (int r = 0; r < 10; r++) { round4(c, k); k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
    }
    __device__ static __forceinline__ double u53(uint32_t hi, uint32_t lo) {
        const uint64_t x = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
        return ((double)x + 1.0)*(1.0/9007199254740992.0);
    }
    // n normals for (seed, stream, step, link): counter (link, step, stream, j),
    // Box-Muller on two 53-bit uniforms.  T = double for the FP64 form (and for
    // the initial conditions of both forms); T = float evaluates the logarithm
    // and the sine/cosine in FP32 from the same uniforms (a tiny u1 keeps its
    // magnitude in FP32, so the tails are those of the FP64 draw; sine and
    // cosine of 2 pi u2 come from sincospif, whose argument reduction is exact)
    // at a fraction of the cost; the FP32 form's noise then differs from the
    // FP64 form's by FP32 rounding.
    template <typename T = double>
    __device__ static __forceinline__ void normals(uint64_t seed, uint32_t stream, uint32_t step, uint32_t link, int n, T* z) {
        for (int j = 0; 2*j < n; j++) {
            uint32_t c[4] = {link, step, stream, (uint32_t)j};
            gen(c, (uint32_t)seed, (uint32_t)(seed >> 32));
            const double u1 = u53(c[0], c[1]), u2 = u53(c[2], c[3]);
            T sn, cs, R;
            if constexpr (sizeof(T) == 8) {
                R = __builtin_sqrt(-2.0*log(u1));
                sincos(6.283185307179586476925286766559*u2, &sn, &cs);
            } else {
                // sin, cos of 2 pi u2 as sin, cos of pi (2 u2): the reduction is exact
                R = __builtin_sqrtf(-2.0f*logf((float)u1));
                sincospif(2.0f*(float)u2, &sn, &cs);
            }
            z[2*j] = R*cs;
            if (2*j + 1 < n) z[2*j + 1] = R*sn;
        }
    }
};

}  // namespace tcym
