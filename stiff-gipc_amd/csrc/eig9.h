// eig9.h -- packed symmetric storage and a register-resident cyclic Jacobi
// eigen-solver for 9x9 matrices (PSD projections of element Hessians).
#pragma once

#include <hip/hip_runtime.h>

namespace sgipc
{

// packed upper-triangle index of (p, q), p <= q < N
template <int N>
__host__ __device__ constexpr int pk(int p, int q)
{
    return p <= q ? p * N - p * (p - 1) / 2 + (q - p) : q * N - q * (q - 1) / 2 + (p - q);
}

// cyclic Jacobi on a packed symmetric 9x9 A; on return A's diagonal holds the
// eigenvalues and V (row-major) the eigenvectors in its columns
__device__ __forceinline__ void jacobi9(double (&A)[45], double (&V)[81])
{
#pragma unroll
    for(int k = 0; k < 81; k++)
        V[k] = (k % 10 == 0) ? 1.0 : 0.0;
    for(int sweep = 0; sweep < 16; sweep++)
    {
        double off = 0.0, dia = 0.0;
#pragma unroll
        for(int p = 0; p < 9; p++)
        {
            dia += A[pk<9>(p, p)] * A[pk<9>(p, p)];
#pragma unroll
            for(int q = p + 1; q < 9; q++)
                off += A[pk<9>(p, q)] * A[pk<9>(p, q)];
        }
        if(!(off > 1e-32 * dia))
            break;
#pragma unroll
        for(int p = 0; p < 8; p++)
#pragma unroll
            for(int q = p + 1; q < 9; q++)
            {
                double apq = A[pk<9>(p, q)];
                if(fabs(apq) > 1e-300)
                {
                    double app = A[pk<9>(p, p)], aqq = A[pk<9>(q, q)];
                    double th  = (aqq - app) / (2.0 * apq);
                    double t   = 1.0 / (fabs(th) + sqrt(th * th + 1.0));
                    if(th < 0.0)
                        t = -t;
                    double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs, tau = sn / (1.0 + cs);
                    A[pk<9>(p, p)] = app - t * apq;
                    A[pk<9>(q, q)] = aqq + t * apq;
                    A[pk<9>(p, q)] = 0.0;
#pragma unroll
                    for(int k = 0; k < 9; k++)
                    {
                        if(k == p || k == q)
                            continue;
                        double akp = A[pk<9>(k, p)], akq = A[pk<9>(k, q)];
                        A[pk<9>(k, p)] = akp - sn * (akq + tau * akp);
                        A[pk<9>(k, q)] = akq + sn * (akp - tau * akq);
                    }
#pragma unroll
                    for(int k = 0; k < 9; k++)
                    {
                        double vkp = V[9 * k + p], vkq = V[9 * k + q];
                        V[9 * k + p] = vkp - sn * (vkq + tau * vkp);
                        V[9 * k + q] = vkq + sn * (vkp - tau * vkq);
                    }
                }
            }
    }
}

}  // namespace sgipc
