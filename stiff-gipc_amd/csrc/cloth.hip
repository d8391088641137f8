// cloth.hip -- cloth membrane (Baraff-Witkin stretch/shear) and discrete-shell
// bending: per-element gradient and PSD Hessian records for the gathers, and
// the line-search energies.
//
//   k_cloth_membrane  _calculate_triangle_fem_gradient_hessian
//                     (femEnergy.cu:2215-2300): F = Ds Dm^-1 (3x2), PK1
//                     (:360-412), stretch Hessian projection (:565-663),
//                     shear projection (:666-740) in closed form
//   k_cloth_bend      _calculate_bending_gradient_hessian (:2431-2566):
//                     H = 2 ((t - t0) d2t + dt dt^T) made PSD (makePDSNK,
//                     :64-84), scaled by k |e0| dt^2
//   k_energy_*        __cal_BaraffWitkinStretch_energy (:1003-1041),
//                     __cal_bending_energy (:2590-2630), both times dt^2
//                     (computeEnergy, GIPC.cu:10411-10416)
//
// Each triangle / hinge writes its record to the element buffer (no atomics);
// the FEM gathers add records into the static BSR blocks and the vertices.
//
// Bending projection: the hinge Hessian is translation invariant, so with
// e = P q (e_k = q_k - q0) and H12 = P^T H9 P, the PSD projection of H12 is
// P^T S^-1 PSD(S H9 S) S^-1 P with S = (P P^T)^(1/2) = (I + J/3) (x) I3 and
// S^-1 = (I - J/6) (x) I3 (J the 3x3 all-ones matrix).  That is a 9x9
// symmetric eigenproblem, solved by cyclic Jacobi in registers, instead of
// the reference's 12x12 one; both give the same projection to rounding.
#include "assembly.h"
#include "cloth.h"
#include "eig9.h"

namespace sgipc
{

// ------------------------------------------------------------- membrane
__global__ void __launch_bounds__(128) k_cloth_membrane(ClothDev c, const double* __restrict__ x,
                                                        double dt2, double* __restrict__ buf,
                                                        int hessian)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if(f >= c.nf)
        return;
    uint32_t t0 = c.tris[3 * (size_t)f], t1 = c.tris[3 * (size_t)f + 1],
             t2 = c.tris[3 * (size_t)f + 2];
    double Dm[4];
#pragma unroll
    for(int k = 0; k < 4; k++)
        Dm[k] = c.dminv[(size_t)k * c.nf + f];
    V3     x0 = ld3(x + 3 * t0);
    V3     d1 = ld3(x + 3 * t1) - x0, d2 = ld3(x + 3 * t2) - x0;
    V3     fu = d1 * Dm[0] + d2 * Dm[2];
    V3     fv = d1 * Dm[1] + d2 * Dm[3];
    double Iu = sqn(fu), Iv = sqn(fv), I6 = dot(fu, fv);
    double su = sqrt(Iu), sv = sqrt(Iv);
    double r  = c.rate;
    // stretch along an axis: f(s) = (s-1)^2 + [s>1] r (s-1)^3; dpsi/dF column
    // = f'(s)/s F column
    auto   d1s = [&](double s, double I) {
        double a = 1.0 - 1.0 / s;
        if(I > 1.0)
            a += 1.5 * r * (s + 1.0 / s - 2.0);
        return 2.0 * a;
    };
    double cu = c.stretch * d1s(su, Iu), cv = c.stretch * d1s(sv, Iv);
    double ch = 2.0 * c.shear * I6;
    V3     Pu = fu * cu + fv * ch;
    V3     Pv = fv * cv + fu * ch;
    double B[3][2] = {{-(Dm[0] + Dm[2]), -(Dm[1] + Dm[3])}, {Dm[0], Dm[1]}, {Dm[2], Dm[3]}};
    double s = c.area[f] * dt2;
    double* ob = buf + c.off_tri + (size_t)kTriStride * f;
#pragma unroll
    for(int a = 0; a < 3; a++)
    {
        V3 g = (Pu * B[a][0] + Pv * B[a][1]) * s;
        ob[54 + 3 * a]     = g.x;
        ob[54 + 3 * a + 1] = g.y;
        ob[54 + 3 * a + 2] = g.z;
    }
    if(!hessian)
        return;

    // stretch: alpha I + beta n n^T per axis (exact for s > 1, projected below)
    double uf[3] = {fu.x, fu.y, fu.z}, vf[3] = {fv.x, fv.y, fv.z};
    double au = 0.0, av = 0.0, bu = 2.0, bv = 2.0;
    if(su > 1.0)
    {
        au = (su - 1.0) * (3.0 * su * r - 3.0 * r + 2.0) / su;
        bu = (3.0 * Iu * r - 3.0 * r + 2.0) / su;
    }
    if(sv > 1.0)
    {
        av = (sv - 1.0) * (3.0 * sv * r - 3.0 * r + 2.0) / sv;
        bv = (3.0 * Iv * r - 3.0 * r + 2.0) / sv;
    }
    // shear: 2 (g g^T + I6 S) projected: eigenvalue lam0 on q0 = I6 S g + lam0 g,
    // |I6| on the sign(I6)-eigenspace of S orthogonal to T q0
    double I2   = Iu + Iv;
    double lam0 = 0.5 * (I2 + sqrt(I2 * I2 + 12.0 * I6 * I6));
    double sg   = I6 >= 0.0 ? 1.0 : -1.0;
    double qu[3], qv[3], tt[3];
    double qn = 0.0;
#pragma unroll
    for(int i = 0; i < 3; i++)
    {
        qu[i] = I6 * uf[i] + lam0 * vf[i];
        qv[i] = I6 * vf[i] + lam0 * uf[i];
        qn += qu[i] * qu[i] + qv[i] * qv[i];
    }
    qn = 1.0 / sqrt(qn);
    double tn = 0.0;
#pragma unroll
    for(int i = 0; i < 3; i++)
    {
        qu[i] *= qn;
        qv[i] *= qn;
        tt[i] = qu[i] + sg * qv[i];
        tn += tt[i] * tt[i];
    }
    tn = 1.0 / tn;
    double aI6 = fabs(I6), l2 = 2.0 * lam0;
    double Huu[9], Huv[9], Hvv[9];
    double isu = 1.0 / Iu, isv = 1.0 / Iv;
#pragma unroll
    for(int i = 0; i < 3; i++)
#pragma unroll
        for(int j = 0; j < 3; j++)
        {
            double P = (i == j ? 1.0 : 0.0) - tt[i] * tt[j] * tn;
            double e = (i == j ? 1.0 : 0.0);
            Huu[3 * i + j] = c.stretch * (au * e + bu * uf[i] * uf[j] * isu)
                             + c.shear * (aI6 * P + l2 * qu[i] * qu[j]);
            Hvv[3 * i + j] = c.stretch * (av * e + bv * vf[i] * vf[j] * isv)
                             + c.shear * (aI6 * P + l2 * qv[i] * qv[j]);
            Huv[3 * i + j] = c.shear * (I6 * P + l2 * qu[i] * qv[j]);
        }
#pragma unroll
    for(int pi = 0; pi < 6; pi++)
    {
        const int a = kTriPairA[pi], b = kTriPairB[pi];
        double    k00 = s * B[a][0] * B[b][0], k01 = s * B[a][0] * B[b][1];
        double    k10 = s * B[a][1] * B[b][0], k11 = s * B[a][1] * B[b][1];
#pragma unroll
        for(int i = 0; i < 3; i++)
#pragma unroll
            for(int j = 0; j < 3; j++)
                ob[9 * pi + 3 * i + j] = k00 * Huu[3 * i + j] + k01 * Huv[3 * i + j]
                                         + k10 * Huv[3 * j + i] + k11 * Hvv[3 * i + j];
    }
}

// ---------------------------------------------------------------- bending
// X <- (I + alpha J) (x) I3 . X . (I + alpha J) (x) I3 on a packed symmetric
// 9x9 whose 3x3 blocks are indexed by edge vector
__device__ __forceinline__ void congruence_j(double (&X)[45], double alpha)
{
    double R[3][9], T[9];
#pragma unroll
    for(int a = 0; a < 3; a++)
#pragma unroll
        for(int i = 0; i < 3; i++)
#pragma unroll
            for(int j = 0; j < 3; j++)
                R[a][3 * i + j] = X[pk<9>(3 * a + i, j)] + X[pk<9>(3 * a + i, 3 + j)]
                                  + X[pk<9>(3 * a + i, 6 + j)];
#pragma unroll
    for(int k = 0; k < 9; k++)
        T[k] = R[0][k] + R[1][k] + R[2][k];
#pragma unroll
    for(int p = 0; p < 9; p++)
#pragma unroll
        for(int q = p; q < 9; q++)
        {
            int a = p / 3, i = p % 3, b = q / 3, j = q % 3;
            X[pk<9>(p, q)] += alpha * (R[a][3 * i + j] + R[b][3 * j + i]) + alpha * alpha * T[3 * i + j];
        }
}

__global__ void __launch_bounds__(64) k_cloth_bend(ClothDev c, const double* __restrict__ x,
                                                   double dt2, double* __restrict__ buf,
                                                   int hessian)
{
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if(h >= c.nh)
        return;
    uint32_t id[4];
#pragma unroll
    for(int k = 0; k < 4; k++)
        id[k] = c.hinges[4 * (size_t)h + k];
    V3     q0 = ld3(x + 3 * id[0]);
    V3     e1 = ld3(x + 3 * id[1]) - q0, e2 = ld3(x + 3 * id[2]) - q0, e3 = ld3(x + 3 * id[3]) - q0;
    V3     c23 = cross(e2, e3), c31 = cross(e3, e1), c12 = cross(e1, e2);
    double V  = dot(e1, c23);
    double s  = sqn(e1), L = sqrt(s);
    double p  = dot(e1, e2), q = dot(e1, e3), t = dot(e2, e3);
    double u  = -V * L, w = p * q - s * t;
    double th = atan2(u, w);
    double r2 = u * u + w * w, ir2 = 1.0 / r2;
    // first derivatives over (e1, e2, e3)
    double E1[3] = {e1.x, e1.y, e1.z}, E2[3] = {e2.x, e2.y, e2.z}, E3[3] = {e3.x, e3.y, e3.z};
    double gV[9] = {c23.x, c23.y, c23.z, c31.x, c31.y, c31.z, c12.x, c12.y, c12.z};
    double gu[9], gw[9], g9[9];
#pragma unroll
    for(int i = 0; i < 3; i++)
    {
        double gL = E1[i] / L;
        gu[i]     = -(L * gV[i] + V * gL);
        gu[3 + i] = -L * gV[3 + i];
        gu[6 + i] = -L * gV[6 + i];
        gw[i]     = q * E2[i] + p * E3[i] - 2.0 * t * E1[i];
        gw[3 + i] = q * E1[i] - s * E3[i];
        gw[6 + i] = p * E1[i] - s * E2[i];
    }
#pragma unroll
    for(int k = 0; k < 9; k++)
        g9[k] = (w * gu[k] - u * gw[k]) * ir2;
    double dl = th - c.theta0[h];
    double sc = c.bend * c.len0[h] * dt2;
    double* ob = buf + c.off_hinge + (size_t)kHingeStride * h;
#pragma unroll
    for(int i = 0; i < 3; i++)
    {
        double gs   = 2.0 * dl * sc;
        ob[90 + i]  = -gs * (g9[i] + g9[3 + i] + g9[6 + i]);
        ob[93 + i]  = gs * g9[i];
        ob[96 + i]  = gs * g9[3 + i];
        ob[99 + i]  = gs * g9[6 + i];
    }
    if(!hessian)
        return;

    // Hessian of theta over e: (w Hu - u Hw)/r2 + [-2uw gu gu^T + 2uw gw gw^T
    // + (u^2 - w^2)(gu gw^T + gw gu^T)]/r2^2, with
    // Hu = -(L HV + gV gL^T + gL gV^T + V HL), HL = (I - e1 e1^T/s)/L on (e1, e1),
    // HV = skew blocks, Hw = q Hp + p Hq - t Hs - s Ht + sym(grad products)
    double A[45];
    double ir4 = ir2 * ir2;
    double cuu = -2.0 * u * w * ir4, cww = 2.0 * u * w * ir4, cuw = (u * u - w * w) * ir4;
    double cu = w * ir2, cw = -u * ir2;
#pragma unroll
    for(int P = 0; P < 9; P++)
#pragma unroll
        for(int Q = P; Q < 9; Q++)
        {
            const int a = P / 3, i = P % 3, b = Q / 3, j = Q % 3;
            // HV block (a, b), a < b: (0,1) -[e3]x, (0,2) [e2]x, (1,2) -[e1]x
            double hv = 0.0;
            if(a != b)
            {
                const double* ev = (a + b == 1) ? E3 : (a + b == 2) ? E2 : E1;
                double        sk = (a + b == 2) ? 1.0 : -1.0;
                double        cx = 0.0;  // ([v]x)_ij
                if(i == 0 && j == 1) cx = -ev[2];
                if(i == 0 && j == 2) cx = ev[1];
                if(i == 1 && j == 0) cx = ev[2];
                if(i == 1 && j == 2) cx = -ev[0];
                if(i == 2 && j == 0) cx = -ev[1];
                if(i == 2 && j == 1) cx = ev[0];
                hv = sk * cx;
            }
            double gLa = (a == 0) ? E1[i] / L : 0.0, gLb = (b == 0) ? E1[j] / L : 0.0;
            double hl  = (a == 0 && b == 0) ? ((i == j ? 1.0 : 0.0) - E1[i] * E1[j] / s) / L : 0.0;
            double hu  = -(L * hv + gV[P] * gLb + gLa * gV[Q] + V * hl);
            // Hw: I blocks (0,1) q, (0,2) p, (0,0) -2t, (1,2) -s, plus products
            double hw = 0.0;
            if(i == j)
            {
                if((a == 0 && b == 1) || (a == 1 && b == 0)) hw += q;
                if((a == 0 && b == 2) || (a == 2 && b == 0)) hw += p;
                if(a == 0 && b == 0) hw -= 2.0 * t;
                if((a == 1 && b == 2) || (a == 2 && b == 1)) hw -= s;
            }
            double gpP = (a == 0) ? E2[i] : (a == 1) ? E1[i] : 0.0;
            double gpQ = (b == 0) ? E2[j] : (b == 1) ? E1[j] : 0.0;
            double gqP = (a == 0) ? E3[i] : (a == 2) ? E1[i] : 0.0;
            double gqQ = (b == 0) ? E3[j] : (b == 2) ? E1[j] : 0.0;
            double gsP = (a == 0) ? 2.0 * E1[i] : 0.0, gsQ = (b == 0) ? 2.0 * E1[j] : 0.0;
            double gtP = (a == 1) ? E3[i] : (a == 2) ? E2[i] : 0.0;
            double gtQ = (b == 1) ? E3[j] : (b == 2) ? E2[j] : 0.0;
            hw += gpP * gqQ + gqP * gpQ - gsP * gtQ - gtP * gsQ;
            double ht = cu * hu + cw * hw + cuu * gu[P] * gu[Q] + cww * gw[P] * gw[Q]
                        + cuw * (gu[P] * gw[Q] + gw[P] * gu[Q]);
            A[pk<9>(P, Q)] = 2.0 * (dl * ht + g9[P] * g9[Q]);
        }
    // PSD projection in the orthonormal coordinates of the translation complement
    double M[45];
#pragma unroll
    for(int k = 0; k < 45; k++)
        M[k] = A[k];
    congruence_j(M, 1.0 / 3.0);
    double Vv[81];
    jacobi9(M, Vv);
    double wmin = M[pk<9>(0, 0)];
#pragma unroll
    for(int k = 1; k < 9; k++)
        wmin = fmin(wmin, M[pk<9>(k, k)]);
    if(wmin < 0.0)
    {
        double lam[9];
#pragma unroll
        for(int k = 0; k < 9; k++)
            lam[k] = fmax(M[pk<9>(k, k)], 0.0);
#pragma unroll
        for(int P = 0; P < 9; P++)
#pragma unroll
            for(int Q = P; Q < 9; Q++)
            {
                double acc = 0.0;
#pragma unroll
                for(int k = 0; k < 9; k++)
                    acc += Vv[9 * P + k] * lam[k] * Vv[9 * Q + k];
                A[pk<9>(P, Q)] = acc;
            }
        congruence_j(A, -1.0 / 6.0);
    }
    // expand to the 4 hinge vertices (q0 = -(e1 + e2 + e3)) and write the 10
    // upper blocks (a <= b, kPairA/kPairB order)
#pragma unroll
    for(int pi = 0; pi < 10; pi++)
    {
        const int va = kPairA[pi], vb = kPairB[pi];
#pragma unroll
        for(int i = 0; i < 3; i++)
#pragma unroll
            for(int j = 0; j < 3; j++)
            {
                double acc = 0.0;
#pragma unroll
                for(int a = 0; a < 3; a++)
#pragma unroll
                    for(int b = 0; b < 3; b++)
                    {
                        double ca = (va == 0) ? -1.0 : (va == a + 1 ? 1.0 : 0.0);
                        double cb = (vb == 0) ? -1.0 : (vb == b + 1 ? 1.0 : 0.0);
                        if(ca != 0.0 && cb != 0.0)
                            acc += ca * cb * A[pk<9>(3 * a + i, 3 * b + j)];
                    }
                ob[9 * pi + 3 * i + j] = sc * acc;
            }
    }
}

// --------------------------------------------------------------- energies
__global__ void __launch_bounds__(256) k_energy_membrane(ClothDev c, const double* __restrict__ x,
                                                         double dt2, double* out)
{
    int    f = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if(f < c.nf)
    {
        uint32_t t0 = c.tris[3 * (size_t)f], t1 = c.tris[3 * (size_t)f + 1],
                 t2 = c.tris[3 * (size_t)f + 2];
        double Dm[4];
        for(int k = 0; k < 4; k++)
            Dm[k] = c.dminv[(size_t)k * c.nf + f];
        V3     x0 = ld3(x + 3 * t0);
        V3     d1 = ld3(x + 3 * t1) - x0, d2 = ld3(x + 3 * t2) - x0;
        V3     fu = d1 * Dm[0] + d2 * Dm[2];
        V3     fv = d1 * Dm[1] + d2 * Dm[3];
        double Iu = sqn(fu), Iv = sqn(fv), I6 = dot(fu, fv);
        double du = sqrt(Iu) - 1.0, dv = sqrt(Iv) - 1.0;
        double st = du * du + (Iu > 1.0 ? c.rate * du * du * du : 0.0) + dv * dv
                    + (Iv > 1.0 ? c.rate * dv * dv * dv : 0.0);
        e = dt2 * c.area[f] * (c.stretch * st + c.shear * I6 * I6);
    }
    block_partial(e, out);
}

__global__ void __launch_bounds__(256) k_energy_bend(ClothDev c, const double* __restrict__ x,
                                                     double dt2, double* out)
{
    int    h = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if(h < c.nh)
    {
        const uint32_t* id = c.hinges + 4 * (size_t)h;
        double d = hinge_theta(x + 3 * id[0], x + 3 * id[1], x + 3 * id[2], x + 3 * id[3])
                   - c.theta0[h];
        e = dt2 * c.bend * d * d * c.len0[h];
    }
    block_partial(e, out);
}

// ----------------------------------------------------------------- launch
void launch_cloth(const SimDev& d, bool hessian, hipStream_t s)
{
    double dt2 = d.dt * d.dt;
    if(d.cloth.nf)
    {
        hipLaunchKernelGGL(k_cloth_membrane, dim3(grid_for(d.cloth.nf, 128)), dim3(128), 0, s,
                           d.cloth, d.x, dt2, d.fem_buf, hessian ? 1 : 0);
        SG_LAUNCH_CHECK();
    }
    if(d.cloth.nh)
    {
        hipLaunchKernelGGL(k_cloth_bend, dim3(grid_for(d.cloth.nh, 64)), dim3(64), 0, s, d.cloth,
                           d.x, dt2, d.fem_buf, hessian ? 1 : 0);
        SG_LAUNCH_CHECK();
    }
}

void launch_energy_cloth(const SimDev& d, const EnergyLayout& L, double* partial, hipStream_t s)
{
    double dt2 = d.dt * d.dt;
    if(L.n[6])
        hipLaunchKernelGGL(k_energy_membrane, dim3(L.n[6]), dim3(256), 0, s, d.cloth, d.x, dt2,
                           partial + L.off[6]);
    if(L.n[7])
        hipLaunchKernelGGL(k_energy_bend, dim3(L.n[7]), dim3(256), 0, s, d.cloth, d.x, dt2,
                           partial + L.off[7]);
    SG_LAUNCH_CHECK();
}

}  // namespace sgipc
