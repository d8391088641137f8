// mas_invert.hip -- MAS bank inverses by scalar Gauss-Jordan in the
// reference's operation order (__inverse6_P96x96, MASPreconditioner.cu:628-724).
//
// Built without FMA contraction: every update is one rounded product and one
// rounded sum, and the pivot division is IEEE-rounded, so for a given bank
// matrix the inverse is the same bit for bit as a sequential restatement of
// the reference's loop.
//
// One wave per bank, lane i < 48 holding column i of the bank matrix in
// registers.  For pivot j the owning lane publishes its column through LDS;
// every lane then divides its row-j entry by the pivot and updates its other
// 47 entries, with column j replaced by e_j first (in place).  The pivot loop
// is unrolled so every register index is static.
//
// Output: the symmetric matrix built from the inverse's upper triangle,
// out[r][c] = Binv[min(r, c)][max(r, c)].  The reference keeps the upper 3x3
// blocks (MasMatrixSymf, eigen_data.h:125-128) after copying each diagonal
// block's upper entries below its diagonal (MASPreconditioner.cu:704-708); the
// operator it applies is exactly this matrix.
#include "linsys.h"

namespace sgipc
{

namespace
{
constexpr int kBD  = 48;  // 3 x 16 DOFs per bank
constexpr int kLd  = 49;  // LDS row pitch of the output staging

template <typename T>
__global__ void __launch_bounds__(64) k_mas_invert_gj(const double* __restrict__ B,
                                                     T* __restrict__ inv)
{
    __shared__ double col[kBD];
    __shared__ double S[kBD * kLd];
    const int     bank = blockIdx.x;
    const int     lane = threadIdx.x;
    const int     i    = lane < kBD ? lane : 0;  // lanes 48..63 shadow column 0
    const double* src  = B + (size_t)bank * kBD * kBD;
    double        c[kBD];
#pragma unroll
    for(int k = 0; k < kBD; k++)
    {
        double v = src[k * kBD + i];
        c[k]     = (k == i && v == 0.0) ? 1.0 : v;
    }
#pragma unroll
    for(int j = 0; j < kBD; j++)
    {
        if(lane == j)
        {
#pragma unroll
            for(int k = 0; k < kBD; k++)
                col[k] = c[k];
        }
        __syncthreads();
        const double rt = col[j];
        const bool   pj = (lane == j);
        const double cj = (pj ? 1.0 : c[j]) / rt;
#pragma unroll
        for(int k = 0; k < kBD; k++)
        {
            if(k == j)
                continue;
            const double rate = -col[k];
            const double t    = rate * cj;
            c[k]              = (pj ? 0.0 : c[k]) + t;
        }
        c[j] = cj;
        __syncthreads();  // col is rewritten by the next pivot's lane
    }
    if(lane < kBD)
    {
#pragma unroll
        for(int k = 0; k < kBD; k++)
            S[k * kLd + lane] = c[k];
    }
    __syncthreads();
    T* dst = inv + (size_t)bank * kBD * kBD;
    for(int e = lane; e < kBD * kBD; e += 64)
    {
        const int r = e / kBD, q = e % kBD;
        const int a = r < q ? r : q, b = r < q ? q : r;
        dst[e]      = (T)S[a * kLd + b];
    }
}
}  // namespace

void launch_mas_invert_gj(const double* B, float* inv, int nbanks, hipStream_t s)
{
    if(nbanks < 1)
        return;
    hipLaunchKernelGGL(k_mas_invert_gj<float>, dim3(nbanks), dim3(64), 0, s, B, inv);
    SG_LAUNCH_CHECK();
}

void launch_mas_invert_gj(const double* B, double* inv, int nbanks, hipStream_t s)
{
    if(nbanks < 1)
        return;
    hipLaunchKernelGGL(k_mas_invert_gj<double>, dim3(nbanks), dim3(64), 0, s, B, inv);
    SG_LAUNCH_CHECK();
}

}  // namespace sgipc
