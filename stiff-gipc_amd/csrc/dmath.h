// dmath.h -- fixed-size FP64 device math for the element kernels.
// Vector ops keep the reference's operation order (gpu_eigen_libs.cu:91-437)
// where a result feeds a comparison (contact classification), so FP64
// outputs match the CPU restatement exactly when built without contraction.
#pragma once

#include <hip/hip_runtime.h>

namespace sgipc
{

struct V3
{
    double x, y, z;
};

__device__ __forceinline__ V3 ld3(const double* p) { return V3{p[0], p[1], p[2]}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double sqn(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// __Determiant of the matrix with columns c0, c1, c2 (gpu_eigen_libs.cu:1193)
__device__ __forceinline__ double det_cols(V3 c0, V3 c1, V3 c2)
{
    double m00 = c0.x, m01 = c1.x, m02 = c2.x;
    double m10 = c0.y, m11 = c1.y, m12 = c2.y;
    double m20 = c0.z, m21 = c1.z, m22 = c2.z;
    return m00 * m11 * m22 + m10 * m21 * m02 + m20 * m01 * m12 - m20 * m11 * m02
           - m00 * m12 * m21 - m01 * m10 * m22;
}

// --------------------------------------------------------- distances
__device__ __forceinline__ double d_pp(V3 v0, V3 v1) { return sqn(v0 - v1); }
__device__ __forceinline__ double d_pt(V3 v0, V3 v1, V3 v2, V3 v3)
{
    V3     b   = cross(v2 - v1, v3 - v1);
    double aTb = dot(v0 - v1, b);
    return aTb * aTb / sqn(b);
}
__device__ __forceinline__ double d_pe(V3 v0, V3 v1, V3 v2)
{
    return sqn(cross(v1 - v0, v2 - v0)) / sqn(v2 - v1);
}
__device__ __forceinline__ double d_ee(V3 v0, V3 v1, V3 v2, V3 v3)
{
    V3     b   = cross(v1 - v0, v3 - v2);
    double aTb = dot(v2 - v0, b);
    return aTb * aTb / sqn(b);
}

// _dType_PT (mlbvh.cu:290-449)
__device__ inline int dtype_pt(V3 v0, V3 v1, V3 v2, V3 v3)
{
    V3 basis0 = v2 - v1, basis1 = v3 - v1, basis2 = v0 - v1;
    V3 nVec   = cross(basis0, basis1);
    basis1    = cross(basis0, nVec);
    double D   = det_cols(basis0, basis1, nVec);
    double p0x = det_cols(basis2, basis1, nVec) / D;
    double p0y = det_cols(basis0, basis2, nVec) / D;
    if(p0x > 0 && p0x < 1 && p0y >= 0)
        return 3;
    basis0     = v3 - v2;
    basis1     = cross(basis0, nVec);
    basis2     = v0 - v2;
    D          = det_cols(basis0, basis1, nVec);
    double p1x = det_cols(basis2, basis1, nVec) / D;
    double p1y = det_cols(basis0, basis2, nVec) / D;
    if(p1x > 0.0 && p1x < 1.0 && p1y >= 0.0)
        return 4;
    basis0     = v1 - v3;
    basis1     = cross(basis0, nVec);
    basis2     = v0 - v3;
    D          = det_cols(basis0, basis1, nVec);
    double p2x = det_cols(basis2, basis1, nVec) / D;
    double p2y = det_cols(basis0, basis2, nVec) / D;
    if(p2x > 0.0 && p2x < 1.0 && p2y >= 0.0)
        return 5;
    if(p0x <= 0.0 && p2x >= 1.0)
        return 0;
    if(p1x <= 0.0 && p0x >= 1.0)
        return 1;
    if(p2x <= 0.0 && p1x >= 1.0)
        return 2;
    return 6;
}

// _dType_EE (mlbvh.cu:451-534)
__device__ inline int dtype_ee(V3 v0, V3 v1, V3 v2, V3 v3)
{
    V3     u = v1 - v0, v = v3 - v2, w = v0 - v2;
    double a = sqn(u), b = dot(u, v), c = sqn(v), d = dot(u, w), e = dot(v, w);
    double D = a * c - b * b, tD = D, sN, tN;
    int    defaultCase = 8;
    sN                 = (b * e - c * d);
    if(sN <= 0.0)
    {
        tN          = e;
        tD          = c;
        defaultCase = 2;
    }
    else if(sN >= D)
    {
        tN          = e + b;
        tD          = c;
        defaultCase = 5;
    }
    else
    {
        tN = (a * e - b * d);
        if(tN > 0.0 && tN < tD && (dot(w, cross(u, v)) == 0.0 || sqn(cross(u, v)) < 1.0e-20 * a * c))
        {
            if(sN < D / 2)
            {
                tN          = e;
                tD          = c;
                defaultCase = 2;
            }
            else
            {
                tN          = e + b;
                tD          = c;
                defaultCase = 5;
            }
        }
    }
    if(tN <= 0.0)
    {
        if(-d <= 0.0)
            return 0;
        else if(-d >= a)
            return 3;
        else
            return 6;
    }
    else if(tN >= tD)
    {
        if((-d + b) <= 0.0)
            return 1;
        else if((-d + b) >= a)
            return 4;
        else
            return 7;
    }
    return defaultCase;
}

__device__ inline double pt_dist_unclassified(V3 p, V3 t0, V3 t1, V3 t2)
{
    switch(dtype_pt(p, t0, t1, t2))
    {
        case 0: return d_pp(p, t0);
        case 1: return d_pp(p, t1);
        case 2: return d_pp(p, t2);
        case 3: return d_pe(p, t0, t1);
        case 4: return d_pe(p, t1, t2);
        case 5: return d_pe(p, t2, t0);
        case 6: return d_pt(p, t0, t1, t2);
        default: return 1e32;
    }
}

__device__ inline double ee_dist_unclassified(V3 a0, V3 a1, V3 b0, V3 b1)
{
    switch(dtype_ee(a0, a1, b0, b1))
    {
        case 0: return d_pp(a0, b0);
        case 1: return d_pp(a0, b1);
        case 2: return d_pe(a0, b0, b1);
        case 3: return d_pp(a1, b0);
        case 4: return d_pp(a1, b1);
        case 5: return d_pe(a1, b0, b1);
        case 6: return d_pe(b0, a0, a1);
        case 7: return d_pe(b1, a0, a1);
        case 8: return d_ee(a0, a1, b0, b1);
        default: return 1e32;
    }
}

// ----------------------------------------------------- 3x3 symmetric eigen
// cyclic Jacobi on a symmetric 3x3 (row-major A); columns of V = vectors
__device__ inline void sym_eig3(double A[9], double w[3], double V[9])
{
    for(int i = 0; i < 9; i++)
        V[i] = (i % 4 == 0) ? 1.0 : 0.0;
    for(int sweep = 0; sweep < 12; sweep++)
    {
        double off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
        double dia = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
        if(off <= 1e-34 * dia || off == 0.0)
            break;
#pragma unroll
        for(int pq = 0; pq < 3; pq++)
        {
            int p = (pq == 2) ? 1 : 0;
            int q = (pq == 0) ? 1 : 2;
            double apq = A[p * 3 + q];
            if(apq == 0.0)
                continue;
            double app = A[p * 3 + p], aqq = A[q * 3 + q];
            double tau = (aqq - app) / (2.0 * apq);
            double t   = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            double c   = 1.0 / sqrt(1.0 + t * t);
            double s   = t * c;
#pragma unroll
            for(int k = 0; k < 3; k++)
            {
                double akp = A[k * 3 + p], akq = A[k * 3 + q];
                A[k * 3 + p] = c * akp - s * akq;
                A[k * 3 + q] = s * akp + c * akq;
            }
#pragma unroll
            for(int k = 0; k < 3; k++)
            {
                double apk = A[p * 3 + k], aqk = A[q * 3 + k];
                A[p * 3 + k] = c * apk - s * aqk;
                A[q * 3 + k] = s * apk + c * aqk;
            }
#pragma unroll
            for(int k = 0; k < 3; k++)
            {
                double vkp = V[k * 3 + p], vkq = V[k * 3 + q];
                V[k * 3 + p] = c * vkp - s * vkq;
                V[k * 3 + q] = s * vkp + c * vkq;
            }
        }
    }
    w[0] = A[0];
    w[1] = A[4];
    w[2] = A[8];
}

// F = U diag(s) V^T with U, V rotations (det +1) for det F > 0.
// V from the eigenvectors of F^T F; U by Gram-Schmidt on F V.
__device__ inline void svd3(const double F[9], double U[9], double s[3], double V[9])
{
    double C[9];
#pragma unroll
    for(int i = 0; i < 3; i++)
#pragma unroll
        for(int j = 0; j < 3; j++)
            C[i * 3 + j] = F[0 * 3 + i] * F[0 * 3 + j] + F[1 * 3 + i] * F[1 * 3 + j]
                           + F[2 * 3 + i] * F[2 * 3 + j];
    double w[3];
    sym_eig3(C, w, V);
    // sort descending
    for(int a = 0; a < 2; a++)
        for(int b = a + 1; b < 3; b++)
            if(w[b] > w[a])
            {
                double t = w[a];
                w[a]     = w[b];
                w[b]     = t;
                for(int k = 0; k < 3; k++)
                {
                    double tv    = V[k * 3 + a];
                    V[k * 3 + a] = V[k * 3 + b];
                    V[k * 3 + b] = tv;
                }
            }
    // det V = +1
    double dV = V[0] * (V[4] * V[8] - V[5] * V[7]) - V[1] * (V[3] * V[8] - V[5] * V[6])
                + V[2] * (V[3] * V[7] - V[4] * V[6]);
    if(dV < 0)
        for(int k = 0; k < 3; k++)
            V[k * 3 + 2] = -V[k * 3 + 2];
    V3 fv[3];
    for(int c = 0; c < 3; c++)
        fv[c] = V3{F[0] * V[c] + F[1] * V[3 + c] + F[2] * V[6 + c],
                   F[3] * V[c] + F[4] * V[3 + c] + F[5] * V[6 + c],
                   F[6] * V[c] + F[7] * V[3 + c] + F[8] * V[6 + c]};
    V3     u0 = fv[0] * (1.0 / sqrt(sqn(fv[0])));
    V3     u1 = fv[1] - u0 * dot(u0, fv[1]);
    u1        = u1 * (1.0 / sqrt(sqn(u1)));
    V3     u2 = cross(u0, u1);
    s[0]      = dot(u0, fv[0]);
    s[1]      = dot(u1, fv[1]);
    s[2]      = dot(u2, fv[2]);
    U[0] = u0.x, U[3] = u0.y, U[6] = u0.z;
    U[1] = u1.x, U[4] = u1.y, U[7] = u1.z;
    U[2] = u2.x, U[5] = u2.y, U[8] = u2.z;
}

}  // namespace sgipc
