// mas_invert.hip -- MAS bank inverses by scalar Gauss-Jordan in the
// reference's operation order (__inverse6_P96x96, MASPreconditioner.cu:628-724).
//
// Built without FMA contraction: every update is one rounded product and one
// rounded sum, and the pivot division is IEEE-rounded, so for a given bank
// matrix the inverse is the same bit for bit as a sequential restatement of
// the reference's loop.
//
// Per pivot j the reference (one thread per column i, matrix a in shared
// memory) does, with every right-hand side read before the step:
//   rt = a[j][j];  colm[k] = a[k][j];  a[k][j] = (k == j);  a[j][i] /= rt;
//   a[k][i] += (-colm[k]) * a[j][i]   for every k != j.
// Each entry therefore sees one fixed sequence of operations, whatever
// thread performs it.  Here one wave eliminates one bank with every entry in
// registers: lane (g, h) = (lane >> 4, lane & 15) holds the 12 x 3 tile of
// rows 12 g .. 12 g + 11 and columns 3 h .. 3 h + 2.  For pivot j the lanes
// owning row j and column j publish them through LDS (double-buffered on the
// parity of j, so one barrier per pivot), then every lane updates its 36
// entries.  The pivot loop is expanded at compile time (template recursion)
// so every register index is static and nothing goes to scratch.
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
constexpr int kBD = 48;  // 3 x 16 DOFs per bank
constexpr int kTR = 12;  // tile rows per lane
constexpr int kTC = 3;   // tile columns per lane
constexpr int kLd = 49;  // LDS row pitch of the output staging

struct GjShared
{
    double row[2][kBD];  // pivot row, by parity of the pivot
    double col[2][kBD];  // pivot column
};

template <int J>
struct GjStep
{
    // one pivot: publish row J and column J, then update the tile
    static __device__ __forceinline__ void run(double (&a)[kTR][kTC], GjShared& sh, int g, int h)
    {
        constexpr int P  = J & 1;
        constexpr int RG = J / kTR, RL = J % kTR;  // lane group and tile row owning row J
        constexpr int CG = J / kTC, CL = J % kTC;  // lane group and tile column owning column J
        if(g == RG)
        {
#pragma unroll
            for(int c = 0; c < kTC; c++)
                sh.row[P][kTC * h + c] = a[RL][c];
        }
        if(h == CG)
        {
#pragma unroll
            for(int r = 0; r < kTR; r++)
                sh.col[P][kTR * g + r] = a[r][CL];
        }
        __syncthreads();
        const double rt = sh.col[P][J];
        double       rj[kTC];
#pragma unroll
        for(int c = 0; c < kTC; c++)
        {
            const bool pc = (kTC * h + c == J);
            rj[c]         = (pc ? 1.0 : sh.row[P][kTC * h + c]) / rt;
        }
#pragma unroll
        for(int r = 0; r < kTR; r++)
        {
            const int row = kTR * g + r;
            if(row == J)
            {
#pragma unroll
                for(int c = 0; c < kTC; c++)
                    a[r][c] = rj[c];
            }
            else
            {
                const double rate = -sh.col[P][row];
#pragma unroll
                for(int c = 0; c < kTC; c++)
                {
                    const bool   pc = (kTC * h + c == J);
                    const double t  = rate * rj[c];
                    a[r][c]         = (pc ? 0.0 : a[r][c]) + t;
                }
            }
        }
        GjStep<J + 1>::run(a, sh, g, h);
    }
};

template <>
struct GjStep<kBD>
{
    static __device__ __forceinline__ void run(double (&)[kTR][kTC], GjShared&, int, int) {}
};

template <typename T>
__global__ void __launch_bounds__(64) k_mas_invert_gj(const double* __restrict__ B,
                                                     T* __restrict__ inv)
{
    __shared__ GjShared sh;
    __shared__ double   S[kBD * kLd];
    const int     bank = blockIdx.x;
    const int     lane = threadIdx.x;
    const int     g = lane >> 4, h = lane & 15;
    const double* src = B + (size_t)bank * kBD * kBD;
    double        a[kTR][kTC];
#pragma unroll
    for(int r = 0; r < kTR; r++)
    {
#pragma unroll
        for(int c = 0; c < kTC; c++)
        {
            const int row = kTR * g + r, col = kTC * h + c;
            double    v   = src[row * kBD + col];
            a[r][c]       = (row == col && v == 0.0) ? 1.0 : v;
        }
    }
    GjStep<0>::run(a, sh, g, h);
#pragma unroll
    for(int r = 0; r < kTR; r++)
    {
#pragma unroll
        for(int c = 0; c < kTC; c++)
            S[(kTR * g + r) * kLd + kTC * h + c] = a[r][c];
    }
    __syncthreads();
    T* dst = inv + (size_t)bank * kBD * kBD;
    for(int e = lane; e < kBD * kBD; e += 64)
    {
        const int r = e / kBD, q = e % kBD;
        const int lo = r < q ? r : q, hi = r < q ? q : r;
        dst[e]       = (T)S[lo * kLd + hi];
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
