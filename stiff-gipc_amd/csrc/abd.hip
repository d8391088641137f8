// abd.hip -- affine-body dynamics on the node system.
//
// Reference: StiffGIPC/abd_system.  A body's state is q = (p, a1, a2, a3);
// its vertices move as x = p + A xbar (A's rows a_k), i.e. x = J q.  The
// kinetic term is carried at vertex level: with x = J q and x~ = J q~,
// sum_i m_i (x_i - x~_i)^2 / 2 = (q - q~)^T M (q - q~) / 2 with
// M = sum_i m_i J_i^T J_i (the reference's dyadic mass), so the body
// gradient M (q - q~) + J^T g_contact is the J^T reduction of the vertex
// gradient, and the body Hessian is M plus the projected shape Hessian plus
// the J^T H J contact blocks the block sink writes.
//
//   k_abd_body    M and PSD(kappa V dt^2 d2 ||A A^T - I||^2) into the 16 node
//                 blocks of each body, shape gradient
//                 (_cal_abd_body_gradient_and_hessian,
//                 setup_abd_system_gradient_and_hessian.cu:196-341;
//                 details/abd_energy.inl:9-127)
//   k_abd_rhs     node RHS: sum_i J_i^T g_i + shape gradient
//                 (_cal_abd_system_barrier_gradient :344-380)
//   k_abd_expand  vertex displacements J dq (cal_dx_from_dq)
//   k_abd_q/x     q = q_s - alpha dq, x = J q (step_forward, cal_x_from_q)
//   k_energy_abd  kappa V dt^2 ||A A^T - I||^2 (cal_abd_shape_energy)
// Fixed bodies keep M and get no shape term and a zero RHS.  Motor bodies add
// 0.5 dq^T P dq with dq = q - q_p on a2, a3 and P = strength M[6:12, 6:12]
// (setup_abd_system_gradient_and_hessian.cu:262-339; the target q_p is set per
// time step on the host).
#include "assembly.h"
#include "eig9.h"

namespace sgipc
{

__global__ void __launch_bounds__(64) k_abd_body(AbdDev a, double dt2, double* __restrict__ vals,
                                                 int hessian)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if(b >= a.nb)
        return;
    const double* q  = a.q + 12 * (size_t)b;
    double*       gb = a.gbody + 12 * (size_t)b;
    bool          fx = a.fixed[b] != 0;
    double        A[3][3];
#pragma unroll
    for(int i = 0; i < 3; i++)
#pragma unroll
        for(int c = 0; c < 3; c++)
            A[i][c] = q[3 + 3 * i + c];
    double D[3][3];  // a_i . a_j - delta_ij
#pragma unroll
    for(int i = 0; i < 3; i++)
#pragma unroll
        for(int j = 0; j < 3; j++)
            D[i][j] = A[i][0] * A[j][0] + A[i][1] * A[j][1] + A[i][2] * A[j][2] - (i == j ? 1.0 : 0.0);
    double kv = a.kappa * a.vol[b] * dt2;
#pragma unroll
    for(int k = 0; k < 3; k++)
        gb[k] = 0.0;
#pragma unroll
    for(int i = 0; i < 3; i++)
#pragma unroll
        for(int c = 0; c < 3; c++)
            gb[3 + 3 * i + c] =
                fx ? 0.0 : kv * 4.0 * (D[i][0] * A[0][c] + D[i][1] * A[1][c] + D[i][2] * A[2][c]);
    // motor: dq on a2, a3 and P dq into the node gradient
    const double* M  = a.M + 144 * (size_t)b;
    const bool    mo = a.motor && a.motor[b] != 0;
    if(mo)
    {
        const double* qm = a.qm + 12 * (size_t)b;
        double        dq[6];
#pragma unroll
        for(int k = 0; k < 6; k++)
            dq[k] = q[6 + k] - qm[6 + k];
#pragma unroll
        for(int r = 0; r < 6; r++)
        {
            double pr = 0.0;
#pragma unroll
            for(int c = 0; c < 6; c++)
                pr += a.motor_strength * M[12 * (6 + r) + 6 + c] * dq[c];
            gb[6 + r] += pr;
        }
    }
    if(!hessian)
        return;
    double S[45];
#pragma unroll
    for(int P = 0; P < 9; P++)
#pragma unroll
        for(int Q = P; Q < 9; Q++)
        {
            int    i = P / 3, c = P % 3, j = Q / 3, d = Q % 3;
            double h;
            if(i == j)
            {
                h = 8.0 * A[i][c] * A[i][d] + 4.0 * D[i][i] * (c == d ? 1.0 : 0.0);
#pragma unroll
                for(int k = 0; k < 3; k++)
                    if(k != i)
                        h += 4.0 * A[k][c] * A[k][d];
            }
            else
                h = 4.0 * A[j][c] * A[i][d] + 4.0 * (D[i][j]) * (c == d ? 1.0 : 0.0);
            S[pk<9>(P, Q)] = fx ? 0.0 : kv * h;
        }
    if(!fx)
    {
        double M9[45], V[81];
#pragma unroll
        for(int k = 0; k < 45; k++)
            M9[k] = S[k];
        jacobi9(M9, V);
        double wmin = M9[pk<9>(0, 0)];
#pragma unroll
        for(int k = 1; k < 9; k++)
            wmin = fmin(wmin, M9[pk<9>(k, k)]);
        if(wmin < 0.0)
        {
            double lam[9];
#pragma unroll
            for(int k = 0; k < 9; k++)
                lam[k] = fmax(M9[pk<9>(k, k)], 0.0);
#pragma unroll
            for(int P = 0; P < 9; P++)
#pragma unroll
                for(int Q = P; Q < 9; Q++)
                {
                    double acc = 0.0;
#pragma unroll
                    for(int k = 0; k < 9; k++)
                        acc += V[9 * P + k] * lam[k] * V[9 * Q + k];
                    S[pk<9>(P, Q)] = acc;
                }
        }
    }
    const int*    sl = a.slots + 16 * (size_t)b;
#pragma unroll
    for(int m = 0; m < 4; m++)
#pragma unroll
        for(int n = 0; n < 4; n++)
        {
            double* dst = vals + 9 * (size_t)sl[4 * m + n];
#pragma unroll
            for(int r = 0; r < 3; r++)
#pragma unroll
                for(int c = 0; c < 3; c++)
                {
                    int    R = 3 * m + r, C = 3 * n + c;
                    double h = M[12 * R + C];
                    if(m > 0 && n > 0)
                        h += S[pk<9>(R - 3, C - 3)];
                    if(mo && m > 1 && n > 1)
                        h += a.motor_strength * M[12 * R + C];
                    dst[3 * r + c] = h;
                }
        }
}

// one 64-lane workgroup per body: sum_i J_i^T (g_contact + g_shape)_i
__global__ void __launch_bounds__(64) k_abd_rhs(AbdDev a, const double* __restrict__ xbar,
                                                const double* __restrict__ gc,
                                                const double* __restrict__ gs,
                                                double* __restrict__ rhs)
{
    int    b = blockIdx.x;
    double acc[12];
#pragma unroll
    for(int k = 0; k < 12; k++)
        acc[k] = 0.0;
    for(int v = a.bstart[b] + (int)threadIdx.x; v < a.bstart[b + 1]; v += 64)
    {
        double f[3], xb[3];
#pragma unroll
        for(int c = 0; c < 3; c++)
        {
            f[c]  = gc[3 * (size_t)v + c] + gs[3 * (size_t)v + c];
            xb[c] = xbar[3 * (size_t)v + c];
        }
#pragma unroll
        for(int k = 0; k < 3; k++)
        {
            acc[k] += f[k];
#pragma unroll
            for(int l = 0; l < 3; l++)
                acc[3 + 3 * k + l] += f[k] * xb[l];
        }
    }
#pragma unroll
    for(int k = 0; k < 12; k++)
        acc[k] = wave_sum(acc[k]);
    if(threadIdx.x == 0)
    {
        double* out = rhs + 3 * ((size_t)a.nfem + 4 * (size_t)b);
        bool    fx  = a.fixed[b] != 0;
#pragma unroll
        for(int k = 0; k < 12; k++)
            out[k] = fx ? 0.0 : acc[k] + a.gbody[12 * (size_t)b + k];
    }
}

// vertex displacements from the node solution: FEM copy, ABD J dq
__global__ void k_abd_expand(int nv, int nfem, const int* __restrict__ node_of,
                             const double* __restrict__ xbar, const double* __restrict__ dy,
                             double* __restrict__ dx)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if(v >= nv)
        return;
    if(v < nfem)
    {
#pragma unroll
        for(int c = 0; c < 3; c++)
            dx[3 * (size_t)v + c] = dy[3 * (size_t)v + c];
        return;
    }
    const double* dq = dy + 3 * (size_t)node_of[v];
    const double* xb = xbar + 3 * (size_t)v;
#pragma unroll
    for(int k = 0; k < 3; k++)
        dx[3 * (size_t)v + k] = dq[k] + dq[3 + 3 * k] * xb[0] + dq[4 + 3 * k] * xb[1]
                                + dq[5 + 3 * k] * xb[2];
}

__global__ void k_abd_q(AbdDev a, const double* __restrict__ dy, double alpha)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= 12 * a.nb)
        return;
    int b = i / 12, k = i % 12;
    if(a.fixed[b])
        return;
    a.q[i] = a.qs[i] - dy[3 * ((size_t)a.nfem + 4 * (size_t)b) + k] * alpha;
}

__global__ void k_abd_x(int nv, int nfem, const int* __restrict__ node_of,
                        const double* __restrict__ xbar, const double* __restrict__ q,
                        double* __restrict__ x)
{
    int v = nfem + blockIdx.x * blockDim.x + threadIdx.x;
    if(v >= nv)
        return;
    const double* qb = q + 12 * (size_t)((node_of[v] - nfem) / 4);
    const double* xb = xbar + 3 * (size_t)v;
#pragma unroll
    for(int k = 0; k < 3; k++)
        x[3 * (size_t)v + k] = qb[k] + qb[3 + 3 * k] * xb[0] + qb[4 + 3 * k] * xb[1]
                               + qb[5 + 3 * k] * xb[2];
}

__global__ void __launch_bounds__(256) k_energy_abd(AbdDev a, double dt2, double* out)
{
    int    b = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if(b < a.nb && !a.fixed[b])
    {
        const double* q = a.q + 12 * (size_t)b;
        for(int i = 0; i < 3; i++)
            for(int j = 0; j < 3; j++)
            {
                double d = q[3 + 3 * i] * q[3 + 3 * j] + q[4 + 3 * i] * q[4 + 3 * j]
                           + q[5 + 3 * i] * q[5 + 3 * j] - (i == j ? 1.0 : 0.0);
                e += d * d;
            }
        e *= a.kappa * a.vol[b] * dt2;
        if(a.motor && a.motor[b])  // the motor's share of the body's kinetic energy
        {
            const double* M  = a.M + 144 * (size_t)b;
            const double* qm = a.qm + 12 * (size_t)b;
            double        dq[6];
            for(int k = 0; k < 6; k++)
                dq[k] = q[6 + k] - qm[6 + k];
            for(int r = 0; r < 6; r++)
            {
                double pr = 0.0;
                for(int c = 0; c < 6; c++)
                    pr += a.motor_strength * M[12 * (6 + r) + 6 + c] * dq[c];
                e += 0.5 * dq[r] * pr;
            }
        }
    }
    block_partial(e, out);
}

// ----------------------------------------------------------------- launch
void launch_abd_body(const SimDev& d, bool hessian, hipStream_t s)
{
    if(!d.abd.nb)
        return;
    hipLaunchKernelGGL(k_abd_body, dim3(grid_for(d.abd.nb, 64)), dim3(64), 0, s, d.abd,
                       d.dt * d.dt, d.vals, hessian ? 1 : 0);
    SG_LAUNCH_CHECK();
}

void launch_abd_rhs(const SimDev& d, double* rhs, hipStream_t s)
{
    if(!d.abd.nb)
        return;
    hipLaunchKernelGGL(k_abd_rhs, dim3(d.abd.nb), dim3(64), 0, s, d.abd, d.xbar, d.gcontact,
                       d.gshape, rhs);
    SG_LAUNCH_CHECK();
}

void launch_abd_expand(const SimDev& d, const double* dy, double* dx, hipStream_t s)
{
    hipLaunchKernelGGL(k_abd_expand, dim3(grid_for(d.nv, 256)), dim3(256), 0, s, d.nv, d.nfem,
                       d.node_of, d.xbar, dy, dx);
    SG_LAUNCH_CHECK();
}

void launch_abd_step(const SimDev& d, const double* dy, double alpha, hipStream_t s)
{
    if(!d.abd.nb)
        return;
    hipLaunchKernelGGL(k_abd_q, dim3(grid_for(12 * d.abd.nb, 256)), dim3(256), 0, s, d.abd, dy,
                       alpha);
    hipLaunchKernelGGL(k_abd_x, dim3(grid_for(d.nv - d.nfem, 256)), dim3(256), 0, s, d.nv,
                       d.nfem, d.node_of, d.xbar, d.abd.q, d.x);
    SG_LAUNCH_CHECK();
}

void launch_energy_abd(const SimDev& d, const EnergyLayout& L, double* partial, hipStream_t s)
{
    if(!L.n[8])
        return;
    hipLaunchKernelGGL(k_energy_abd, dim3(L.n[8]), dim3(256), 0, s, d.abd, d.dt * d.dt,
                       partial + L.off[8]);
    SG_LAUNCH_CHECK();
}

}  // namespace sgipc
