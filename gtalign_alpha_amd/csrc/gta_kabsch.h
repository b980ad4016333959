// gta_kabsch.h -- the 3x3 superposition solve on the device.
//
// Kabsch's closed-form method as the reference applies it
// (transformbase.h:167-583, MpStageBase.h CalcTfmMatricesHelper_Complete):
// centre, R = sum(q r^T) - n qbar rbar^T scaled by the mean |R_ij|, the
// eigenvalues of R^T R from the trigonometric solution of the characteristic
// cubic, eigenvectors, rotation, translation, with the same degeneracy
// thresholds and identity fallback.  T = float reproduces the reference's
// single-precision solve; T = double is the mixed-precision form's FP64 solve.
// Every lane of a wave evaluates the solve on identical (uniform) inputs.
#pragma once

#include "gta_device.h"

namespace gta {

template <typename T> __device__ __forceinline__ T t_sqrt(T x);
template <> __device__ __forceinline__ float t_sqrt<float>(float x) { return sqrtf(x); }
template <> __device__ __forceinline__ double t_sqrt<double>(double x) { return sqrt(x); }
template <typename T> __device__ __forceinline__ T t_abs(T x) { return x < (T)0 ? -x : x; }
template <typename T> __device__ __forceinline__ T t_atan2(T y, T x);
template <> __device__ __forceinline__ float t_atan2<float>(float y, float x) { return atan2f(y, x); }
template <> __device__ __forceinline__ double t_atan2<double>(double y, double x) { return atan2(y, x); }
template <typename T> __device__ __forceinline__ void t_sincos(T a, T& s, T& c);
template <> __device__ __forceinline__ void t_sincos<float>(float a, float& s, float& c) { s = sinf(a); c = cosf(a); }
template <> __device__ __forceinline__ void t_sincos<double>(double a, double& s, double& c) { s = sin(a); c = cos(a); }

template <typename T>
__device__ __forceinline__ void CalcRmatrix(T* c) {
    T c01 = c[1], c02 = c[2], c12 = c[5];
    c[0] = c[0] - c[CVq0 + 0] * c[CVr0 + 0];
    c[1] = c[3] - c[CVq0 + 1] * c[CVr0 + 0];
    c[2] = c[6] - c[CVq0 + 2] * c[CVr0 + 0];
    c[3] = c01 - c[CVq0 + 0] * c[CVr0 + 1];
    c[4] = c[4] - c[CVq0 + 1] * c[CVr0 + 1];
    c[5] = c[7] - c[CVq0 + 2] * c[CVr0 + 1];
    c[6] = c02 - c[CVq0 + 0] * c[CVr0 + 2];
    c[7] = c12 - c[CVq0 + 1] * c[CVr0 + 2];
    c[8] = c[8] - c[CVq0 + 2] * c[CVr0 + 2];
}
template <typename T>
__device__ __forceinline__ T CalcDet(const T* c) {
    return c[0] * (c[4] * c[8] - c[5] * c[7]) - c[1] * (c[3] * c[8] - c[5] * c[6]) +
           c[2] * (c[3] * c[7] - c[4] * c[6]);
}
template <typename T>
__device__ __forceinline__ void CalcRTR(const T* c, T* rr) {
    rr[0] = c[0] * c[0] + c[3] * c[3] + c[6] * c[6];
    rr[1] = c[0] * c[1] + c[3] * c[4] + c[6] * c[7];
    rr[2] = c[1] * c[1] + c[4] * c[4] + c[7] * c[7];
    rr[3] = c[0] * c[2] + c[3] * c[5] + c[6] * c[8];
    rr[4] = c[1] * c[2] + c[4] * c[5] + c[7] * c[8];
    rr[5] = c[2] * c[2] + c[5] * c[5] + c[8] * c[8];
}
template <typename T>
__device__ __forceinline__ bool SolveCubic(T det, T spur, T cof, T& e0, T& e1, T& e2) {
    T h = spur * spur - cof;
    if (h <= (T)0) return false;
    T g = (spur * cof - det * det) * (T)0.5 - spur * h;
    T sqrth = t_sqrt(h);
    T hh = h * h * h - g * g;
    T d = t_atan2(t_sqrt(hh > (T)0 ? hh : (T)0), -g) * ((T)1 / (T)3);
    T sth, cth;
    t_sincos(d, sth, cth);
    cth = sqrth * cth;
    sth = sqrth * (T)1.7320508075688772 * sth;
    e0 = (spur + cth) + cth;
    e1 = (spur - cth) + sth;
    e2 = (spur - cth) - sth;
    return true;
}
template <typename T>
__device__ __forceinline__ void CalcPartialA(int col, T d, const T* rr, T* a) {
    const T eps = (T)TFMEPSILON;
    T ss0 = (d - rr[2]) * (d - rr[5]) - rr[4] * rr[4];
    T ss1 = (d - rr[5]) * rr[1] + rr[3] * rr[4];
    T ss2 = (d - rr[0]) * (d - rr[5]) - rr[3] * rr[3];
    T ss3 = (d - rr[2]) * rr[3] + rr[1] * rr[4];
    T ss4 = (d - rr[0]) * rr[4] + rr[1] * rr[3];
    T ss5 = (d - rr[0]) * (d - rr[2]) - rr[1] * rr[1];
    if (t_abs(ss0) <= eps) ss0 = 0;
    if (t_abs(ss1) <= eps) ss1 = 0;
    if (t_abs(ss2) <= eps) ss2 = 0;
    if (t_abs(ss3) <= eps) ss3 = 0;
    if (t_abs(ss4) <= eps) ss4 = 0;
    if (t_abs(ss5) <= eps) ss5 = 0;
    T v0, v1, v2;
    if (t_abs(ss0) >= t_abs(ss2)) {
        if (t_abs(ss0) < t_abs(ss5)) { v0 = ss3; v1 = ss4; v2 = ss5; }
        else { v0 = ss0; v1 = ss1; v2 = ss3; }
    } else if (t_abs(ss2) >= t_abs(ss5)) {
        v0 = ss1; v1 = ss2; v2 = ss4;
    } else {
        v0 = ss3; v1 = ss4; v2 = ss5;
    }
    T n = v0 * v0 + v1 * v1 + v2 * v2;
    n = (n > eps) ? (T)1 / t_sqrt(n) : (T)0;
    a[col] = v0 * n; a[col + 3] = v1 * n; a[col + 6] = v2 * n;
}
// Completion of a unit vector x orthogonal to y when x has degenerated
// (transformbase.h:400-420): the row r of y with the smallest magnitude
// (last one on ties, starting from 1) gets x_r = 0, and the other two rows
// k = r+1, l = r+2 (mod 3) get x_k = -y_l/p, x_l = y_k/p, p = |(y_k, y_l)|.
// Written with selects so that no array is indexed at run time.
template <typename T>
__device__ __forceinline__ bool complete_orthogonal(T y0, T y1, T y2, T& x0, T& x1, T& x2) {
    T p = 1;
    int r = 2;
    if (t_abs(y0) <= p) { p = t_abs(y0); r = 0; }
    if (t_abs(y1) <= p) { p = t_abs(y1); r = 1; }
    if (t_abs(y2) <= p) { p = t_abs(y2); r = 2; }
    const T yk = (r == 0) ? y1 : ((r == 1) ? y2 : y0);
    const T yl = (r == 0) ? y2 : ((r == 1) ? y0 : y1);
    p = t_sqrt(yk * yk + yl * yl);
    if (p <= (T)TFMTOL) return false;
    const T xk = -yl / p, xl = yk / p;
    x0 = (r == 0) ? (T)0 : ((r == 1) ? xl : xk);
    x1 = (r == 1) ? (T)0 : ((r == 2) ? xl : xk);
    x2 = (r == 2) ? (T)0 : ((r == 0) ? xl : xk);
    return true;
}

template <typename T>
__device__ __forceinline__ bool CalcCompleteA(T e0, T e1, T e2, T* a) {
    T d = a[0] * a[2] + a[3] * a[5] + a[6] * a[8];
    // x: the column to orthogonalise (m1), y: the other eigenvector column (m)
    const bool sw = e0 - e1 > e1 - e2;   // m1 = 2, m = 0; otherwise m1 = 0, m = 2
    const T y0 = sw ? a[0] : a[2], y1 = sw ? a[3] : a[5], y2 = sw ? a[6] : a[8];
    T x0 = sw ? a[2] : a[0], x1 = sw ? a[5] : a[3], x2 = sw ? a[8] : a[6];
    T p = 0;
    x0 -= d * y0; p += x0 * x0;
    x1 -= d * y1; p += x1 * x1;
    x2 -= d * y2; p += x2 * x2;
    if (p > (T)TFMTOL) {
        p = (T)1 / t_sqrt(p);
        x0 *= p; x1 *= p; x2 *= p;
    } else if (!complete_orthogonal(y0, y1, y2, x0, x1, x2)) {
        return false;
    }
    if (sw) { a[2] = x0; a[5] = x1; a[8] = x2; }
    else { a[0] = x0; a[3] = x1; a[6] = x2; }
    a[1] = a[5] * a[6] - a[3] * a[8];
    a[4] = a[8] * a[0] - a[6] * a[2];
    a[7] = a[2] * a[3] - a[0] * a[5];
    return true;
}
template <typename T>
__device__ __forceinline__ bool CalcRotMtx(const T* a, T* b) {
    const T eps = (T)TFMEPSILON;
    T d = 0;
    T r0 = b[0], r1 = b[3], r2 = b[6];
    b[0] = r0 * a[0] + b[1] * a[3] + b[2] * a[6]; d += b[0] * b[0];
    b[3] = r1 * a[0] + b[4] * a[3] + b[5] * a[6]; d += b[3] * b[3];
    b[6] = r2 * a[0] + b[7] * a[3] + b[8] * a[6]; d += b[6] * b[6];
    d = (d > eps) ? (T)1 / t_sqrt(d) : (T)0;
    b[0] *= d; b[3] *= d; b[6] *= d;
    d = 0;
    b[1] = r0 * a[1] + b[1] * a[4] + b[2] * a[7]; d += b[1] * b[1];
    b[4] = r1 * a[1] + b[4] * a[4] + b[5] * a[7]; d += b[4] * b[4];
    b[7] = r2 * a[1] + b[7] * a[4] + b[8] * a[7]; d += b[7] * b[7];
    d = (d > eps) ? (T)1 / t_sqrt(d) : (T)0;
    b[1] *= d; b[4] *= d; b[7] *= d;
    d = b[0] * b[1] + b[3] * b[4] + b[6] * b[7];
    T p = 0;
    b[1] -= d * b[0]; p += b[1] * b[1];
    b[4] -= d * b[3]; p += b[4] * b[4];
    b[7] -= d * b[6]; p += b[7] * b[7];
    if (p > (T)TFMTOL) {
        p = (T)1 / t_sqrt(p);
        b[1] *= p; b[4] *= p; b[7] *= p;
    } else {
        T x0, x1, x2;
        if (!complete_orthogonal(b[0], b[3], b[6], x0, x1, x2)) return false;
        b[1] = x0; b[4] = x1; b[7] = x2;
    }
    b[2] = b[3] * b[7] - b[4] * b[6];
    b[5] = b[6] * b[1] - b[7] * b[0];
    b[8] = b[0] * b[4] - b[1] * b[3];
    r0 = b[0]; r1 = b[1]; r2 = b[2];
    b[0] = r0 * a[0] + r1 * a[1] + r2 * a[2];
    b[1] = r0 * a[3] + r1 * a[4] + r2 * a[5];
    b[2] = r0 * a[6] + r1 * a[7] + r2 * a[8];
    r0 = b[3]; r1 = b[4]; r2 = b[5];
    b[3] = r0 * a[0] + r1 * a[1] + r2 * a[2];
    b[4] = r0 * a[3] + r1 * a[4] + r2 * a[5];
    b[5] = r0 * a[6] + r1 * a[7] + r2 * a[8];
    r0 = b[6]; r1 = b[7]; r2 = b[8];
    b[6] = r0 * a[0] + r1 * a[1] + r2 * a[2];
    b[7] = r0 * a[3] + r1 * a[4] + r2 * a[5];
    b[8] = r0 * a[6] + r1 * a[7] + r2 * a[8];
    return true;
}

// cc: 15 sums (cross terms, query sum, reference sum) and n; the result is the
// 12-float transform query -> reference.  dinv swaps the roles when the query
// is the shorter structure and inverts the result (MpStageBase.h:751-777).
template <typename T>
__device__ __forceinline__ void kabsch(const T* ccin, T n, bool dinv, float* tfm) {
    T c[15];
#pragma unroll
    for (int i = 0; i < 15; i++) c[i] = ccin[i];
    if (dinv) {
        T t;
        t = c[1]; c[1] = c[3]; c[3] = t;
        t = c[2]; c[2] = c[6]; c[6] = t;
        t = c[5]; c[5] = c[7]; c[7] = t;
        for (int i = 0; i < 3; i++) { t = c[CVq0 + i]; c[CVq0 + i] = c[CVr0 + i]; c[CVr0 + i] = t; }
    }
    T a[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, rr[6];
    for (int i = CVq0; i < CVq0 + 3; i++) c[i] /= n;
    CalcRmatrix(c);
    for (int i = CVr0; i < CVr0 + 3; i++) c[i] /= n;
    T scale = (t_abs(c[0]) + t_abs(c[1]) + t_abs(c[2]) + t_abs(c[3]) + t_abs(c[4]) + t_abs(c[5]) +
               t_abs(c[6]) + t_abs(c[7]) + t_abs(c[8])) * (T)(1.0f / 9.0f);
    for (int i = 0; i < 9; i++) c[i] /= scale;
    T det = CalcDet(c);
    CalcRTR(c, rr);
    const T third = (T)(1.0f / 3.0f);
    T spur = (rr[0] + rr[2] + rr[5]) * third;
    T cof = (((((rr[2] * rr[5] - rr[4] * rr[4]) + rr[0] * rr[5]) - rr[3] * rr[3]) + rr[0] * rr[2]) -
             rr[1] * rr[1]) * third;
    bool ok = (spur > (T)0);
    if (ok) {
        T e0, e1, e2;
        if (SolveCubic(det, spur, cof, e0, e1, e2)) {
            CalcPartialA((int)0, e0, rr, a);
            CalcPartialA((int)2, e2, rr, a);
            ok = CalcCompleteA(e0, e1, e2, a);
        }
        if (ok) ok = CalcRotMtx(a, c);
    }
    if (!ok) {
        for (int i = 0; i < 9; i++) c[i] = 0;
        c[0] = c[4] = c[8] = 1;
    }
    T x0 = c[CVq0], x1 = c[CVq0 + 1], x2 = c[CVq0 + 2];
    T t0 = c[CVr0 + 0] - c[0] * x0 - c[1] * x1 - c[2] * x2;
    T t1 = c[CVr0 + 1] - c[3] * x0 - c[4] * x1 - c[5] * x2;
    T t2 = c[CVr0 + 2] - c[6] * x0 - c[7] * x1 - c[8] * x2;
    if (dinv) {
        T t;
        t = c[1]; c[1] = c[3]; c[3] = t;
        t = c[2]; c[2] = c[6]; c[6] = t;
        t = c[5]; c[5] = c[7]; c[7] = t;
        T y0 = c[0] * t0 + c[1] * t1 + c[2] * t2;
        T y1 = c[3] * t0 + c[4] * t1 + c[5] * t2;
        T y2 = c[6] * t0 + c[7] * t1 + c[8] * t2;
        t0 = -y0; t1 = -y1; t2 = -y2;
    }
    for (int i = 0; i < 9; i++) tfm[i] = (float)c[i];
    tfm[9] = (float)t0; tfm[10] = (float)t1; tfm[11] = (float)t2;
}

// RMSD from Kabsch's E0 identity (MpStageFin.h:477-540); cc as above plus the
// six sums of squares sq[0..5] (query x,y,z then reference x,y,z)
template <typename T>
__device__ __forceinline__ float rmsd_e0(const T* ccin, const T* sq, T n) {
    if (n <= (T)0) return 0.f;
    T c[15], s2[6], rr[6];
#pragma unroll
    for (int i = 0; i < 15; i++) c[i] = ccin[i];
    for (int i = CVq0; i < CVq0 + 3; i++) c[i] /= n;
    for (int i = 0; i < 3; i++) s2[i] = sq[i] / n;
    CalcRmatrix(c);
    for (int i = CVr0; i < CVr0 + 3; i++) c[i] /= n;
    for (int i = 3; i < 6; i++) s2[i] = sq[i] / n;
    T scale = (t_abs(c[0]) + t_abs(c[1]) + t_abs(c[2]) + t_abs(c[3]) + t_abs(c[4]) + t_abs(c[5]) +
               t_abs(c[6]) + t_abs(c[7]) + t_abs(c[8])) * (T)(1.0f / 9.0f);
    for (int i = 0; i < 9; i++) c[i] /= scale;
    T det = CalcDet(c);
    CalcRTR(c, rr);
    T E0 = s2[0] - c[CVq0] * c[CVq0] + s2[1] - c[CVq0 + 1] * c[CVq0 + 1] + s2[2] -
           c[CVq0 + 2] * c[CVq0 + 2] + s2[3] - c[CVr0] * c[CVr0] + s2[4] - c[CVr0 + 1] * c[CVr0 + 1] +
           s2[5] - c[CVr0 + 2] * c[CVr0 + 2];
    const T third = (T)(1.0f / 3.0f);
    T spur = (rr[0] + rr[2] + rr[5]) * third;
    T cof = (((((rr[2] * rr[5] - rr[4] * rr[4]) + rr[0] * rr[5]) - rr[3] * rr[3]) + rr[0] * rr[2]) -
             rr[1] * rr[1]) * third;
    T e0 = spur, e1 = spur, e2 = spur;
    if (spur > (T)0) SolveCubic(det, spur, cof, e0, e1, e2);
    e0 = (e0 <= (T)0) ? (T)0 : t_sqrt(e0);
    e1 = (e1 <= (T)0) ? (T)0 : t_sqrt(e1);
    e2 = (e2 <= (T)0) ? (T)0 : (t_sqrt(e2) * ((det < (T)0) ? (T)-1 : (T)1));
    E0 -= ((T)2 * scale * (e0 + e1 + e2)) / n;
    return (float)((E0 <= (T)0) ? (T)0 : t_sqrt(E0));
}

}  // namespace gta
