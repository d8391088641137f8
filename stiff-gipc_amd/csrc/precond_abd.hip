// precond_abd.hip -- block-Jacobi preconditioning of affine bodies: the 12x12
// diagonal block of each body (its 4 nodes p, a1, a2, a3) inverted once per
// Newton iteration and applied in the PCG (ABDPreconditioner,
// linear_system/preconditioner/abd_preconditioner.cu:15-38, with its set-up
// ABDSystem::_cal_abd_system_preconditioner, abd_system/abd_system_function/
// setup_abd_system_gradient_and_hessian.cu:603-646).
//
// The reference inverts each 12x12 with Eigen's inverse(); here Gauss-Jordan
// with partial pivoting (first row of largest |pivot|, multiplication by the
// pivot's reciprocal, rows with a zero multiplier skipped), the elimination
// the oracle restates.  Built without FMA contraction, so both give the same
// inverse bit for bit.
#include "linsys.h"

#include "../../include/sgipc.h"

namespace sgipc
{

namespace
{
constexpr int kB = 12;       // DOFs per body
constexpr int kW = 2 * kB;   // augmented row [A | I]

// one 64-lane workgroup per body; [A | I] in LDS
__global__ void __launch_bounds__(64) k_body_inverse(const int* __restrict__ slots,
                                                     const double* __restrict__ vals,
                                                     double* __restrict__ binv)
{
    __shared__ double M[kB][kW];
    __shared__ int    piv;
    __shared__ int    singular;
    const int b = blockIdx.x, lane = threadIdx.x;
    // gather the 16 node blocks (row-major 3x3 at BSR slot slots[16 b + 4 m + n])
    for(int e = lane; e < kB * kB; e += 64)
    {
        const int row = e / kB, col = e % kB;
        const int m = row / 3, n = col / 3;
        const int s = slots[16 * (size_t)b + 4 * m + n];
        M[row][col]     = vals[9 * (size_t)s + 3 * (row % 3) + (col % 3)];
        M[row][kB + col] = (row == col) ? 1.0 : 0.0;
    }
    if(lane == 0)
        singular = 0;
    __syncthreads();
    for(int c = 0; c < kB; c++)
    {
        if(lane == 0)
        {
            int    p  = c;
            double mx = fabs(M[c][c]);
            for(int r = c + 1; r < kB; r++)
                if(fabs(M[r][c]) > mx)
                {
                    mx = fabs(M[r][c]);
                    p  = r;
                }
            piv = p;
            if(mx == 0.0)
                singular = 1;
        }
        __syncthreads();
        if(singular)
            break;
        const int p = piv;
        if(p != c && lane < kW)
        {
            double t    = M[c][lane];
            M[c][lane]  = M[p][lane];
            M[p][lane]  = t;
        }
        __syncthreads();
        const double inv = 1.0 / M[c][c];
        __syncthreads();
        if(lane < kW)
            M[c][lane] *= inv;
        __syncthreads();
        // rows r != c: M[r][k] -= f M[c][k] with f = M[r][c] taken before the update
        double f[kB];
#pragma unroll
        for(int r = 0; r < kB; r++)
            f[r] = M[r][c];
        __syncthreads();
        for(int e = lane; e < kB * kW; e += 64)
        {
            const int r = e / kW, k = e % kW;
            if(r != c && f[r] != 0.0)
                M[r][k] -= f[r] * M[c][k];
        }
        __syncthreads();
    }
    double* out = binv + (size_t)b * kB * kB;
    for(int e = lane; e < kB * kB; e += 64)
    {
        const int row = e / kB, col = e % kB;
        // a singular block (a body without mass) is left as the identity
        out[e] = singular ? (row == col ? 1.0 : 0.0) : M[row][kB + col];
    }
}

// optional update x += a p, r -= a Ap on the body's 12 DOFs, then
// z = Binv r and r.z into the slot set rz_out
template <typename V>
__global__ void __launch_bounds__(256)
    k_update_body(int nbody, int first, V* x, V* r, const V* p, const V* Ap,
                  const double* rz_prev, const double* pap, const V* __restrict__ binv, V* z,
                  double* rz_out, const int* __restrict__ done)
{
    if(done && *done)
        return;
    double a = 0.0;
    if(x)
        a = slot_set_sum(rz_prev) / slot_set_sum(pap);
    const V av = (V)a;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.0;
    if(b < nbody)
    {
        const size_t o = 3 * ((size_t)first + 4 * (size_t)b);
        V            rb[kB];
        for(int i = 0; i < kB; i++)
        {
            rb[i] = r[o + i];
            if(x)
            {
                x[o + i] += av * p[o + i];
                rb[i] -= av * Ap[o + i];
                r[o + i] = rb[i];
            }
        }
        const V* B = binv + (size_t)b * kB * kB;
        for(int i = 0; i < kB; i++)
        {
            V zi = 0;
            for(int k = 0; k < kB; k++)
                zi += B[i * kB + k] * rb[k];
            z[o + i] = zi;
            d += (double)(rb[i] * zi);
        }
    }
    if(rz_out)
        slot_set_add(rz_out, d);
}
}  // namespace

void launch_body_inverse(int nbody, const int* slots, const double* vals, double* binv,
                         hipStream_t s)
{
    if(nbody < 1)
        return;
    hipLaunchKernelGGL(k_body_inverse, dim3(nbody), dim3(64), 0, s, slots, vals, binv);
    SG_LAUNCH_CHECK();
}

template <typename V>
void launch_update_body(int nbody, int first, V* x, V* r, const V* p, const V* Ap,
                        const double* rz_prev, const double* pap, const V* binv, V* z,
                        double* rz_out, const int* done, hipStream_t s)
{
    if(nbody < 1)
        return;
    SG_CHECK(grid_for(nbody, 256) <= (unsigned)kSetMaxBlocks, SGIPC_ECAPACITY,
             "too many affine bodies for the fixed-order reductions");
    hipLaunchKernelGGL(k_update_body<V>, dim3(grid_for(nbody, 256)), dim3(256), 0, s, nbody,
                       first, x, r, p, Ap, rz_prev, pap, binv, z, rz_out, done);
    SG_LAUNCH_CHECK();
    if(rz_out)
        set_finalize(rz_out, (int)grid_for(nbody, 256), done, s);
}

template void launch_update_body<double>(int, int, double*, double*, const double*,
                                         const double*, const double*, const double*,
                                         const double*, double*, double*, const int*,
                                         hipStream_t);
template void launch_update_body<float>(int, int, float*, float*, const float*, const float*,
                                        const double*, const double*, const float*, float*,
                                        double*, const int*, hipStream_t);

}  // namespace sgipc
