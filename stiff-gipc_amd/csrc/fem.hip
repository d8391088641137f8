// fem.hip -- gradient/Hessian assembly and incremental-potential energies.
//
//   k_init_diag   kinetic gradient m(x - x~) and the m I diagonal blocks
//                 (_calKineticGradient GIPC.cu:6341, mass triplets :10201-10220)
//   k_fem         SNK1 tet gradient and PSD Hessian (_calculate_fem_gradient_hessian,
//                 femEnergy.cu:1789-1926; PK1 :443-486; Hessian :1706-1786),
//                 written per tet to a buffer; k_fem_gather_* add the 10 upper
//                 blocks into the static BSR (both orientations) and the
//                 gradient into the vertices without atomics
//   k_ground_*    ground IPC barrier (GIPC.cu:6435-6516)
//   k_barrier     GIPC rank-2 barrier, gradient b'(s) ds/dx and Hessian
//                 lambda0 ds/dx ds/dx^T, s = d/sqrt(dhat^2)
//                 (_calBarrierGradientAndHessian GIPC.cu:3108-4947)
//   k_energy_*    FP64 energies for the line search (computeEnergy :10392-10451)
//                 with fixed-order two-pass reductions (run-to-run identical).
//
// Layout: per-tet constants are SoA (dminv[k * nt + e]) so each of the nine
// loads is a coalesced 512-byte wave access.  FP64 atomics from every tet
// into shared BSR blocks are avoided: each BSR block (and vertex) owns a
// precomputed list of the tet blocks that add into it.
#include <climits>

#include "assembly.h"

namespace sgipc
{

// kinetic gradient of every vertex (affine-body vertices included: their
// J^T reduction is M (q - q~)); m I blocks for the FEM vertices
__global__ void k_init_diag(int nv, int nfem, const double* __restrict__ mass,
                            const double* __restrict__ x, const double* __restrict__ xt,
                            const int* __restrict__ diag, double* __restrict__ vals,
                            double* __restrict__ gshape, double* __restrict__ gcontact)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= nv)
        return;
    double m = mass[i];
    if(i < nfem)
    {
        double* B = vals + 9 * (size_t)diag[i];
        B[0] = m, B[1] = 0, B[2] = 0, B[3] = 0, B[4] = m, B[5] = 0, B[6] = 0, B[7] = 0, B[8] = m;
    }
    for(int c = 0; c < 3; c++)
    {
        gshape[3 * i + c]   = (x[3 * i + c] - xt[3 * i + c]) * m;
        gcontact[3 * i + c] = 0.0;
    }
}

// animated soft targets (_computeSoftConstraintGradientAndHessian,
// GIPC.cu:6365-6409): gradient d rate rate (x - t) added to the vertex's
// shape gradient and rate rate d on its diagonal block (each vertex carries at
// most one target, so no two lanes touch one vertex)
__global__ void k_soft(int n, const uint32_t* __restrict__ idx, const double* __restrict__ tgt,
                       const double* __restrict__ x, double d, double rate,
                       const int* __restrict__ diag, double* __restrict__ vals,
                       double* __restrict__ gshape, int hessian)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n)
        return;
    const size_t v  = idx[i];
    const double gk = d * rate * rate;
    for(int c = 0; c < 3; c++)
        gshape[3 * v + c] += gk * (x[3 * v + c] - tgt[3 * (size_t)i + c]);
    if(hessian)
    {
        const double hk = rate * rate * d;
        double*      B  = vals + 9 * (size_t)diag[v];
        B[0] += hk;
        B[4] += hk;
        B[8] += hk;
    }
}

// upper-triangle index of (p, q), p <= q < 12
__device__ __forceinline__ int uidx(int p, int q) { return p * 12 - p * (p - 1) / 2 + (q - p); }

__global__ void __launch_bounds__(128) k_fem(int nt, const uint32_t* __restrict__ tets,
                                             const double* __restrict__ dminv,
                                             const double* __restrict__ vol,
                                             const double* __restrict__ mu_,
                                             const double* __restrict__ lam_,
                                             const double* __restrict__ x, double dt2,
                                             double* __restrict__ out, int hessian)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if(e >= nt)
        return;
    uint32_t t[4] = {tets[4 * (size_t)e], tets[4 * (size_t)e + 1], tets[4 * (size_t)e + 2],
                     tets[4 * (size_t)e + 3]};
    double   Dm[9];
#pragma unroll
    for(int k = 0; k < 9; k++)
        Dm[k] = dminv[(size_t)k * nt + e];
    V3     x0 = ld3(x + 3 * t[0]);
    V3     d1 = ld3(x + 3 * t[1]) - x0, d2 = ld3(x + 3 * t[2]) - x0, d3 = ld3(x + 3 * t[3]) - x0;
    double Ds[9] = {d1.x, d2.x, d3.x, d1.y, d2.y, d3.y, d1.z, d2.z, d3.z};
    double F[9];
#pragma unroll
    for(int i = 0; i < 3; i++)
#pragma unroll
        for(int j = 0; j < 3; j++)
            F[i * 3 + j] = Ds[i * 3 + 0] * Dm[0 * 3 + j] + Ds[i * 3 + 1] * Dm[1 * 3 + j]
                           + Ds[i * 3 + 2] * Dm[2 * 3 + j];
    double mu = mu_[e], lam = lam_[e];
    // cofactors and J
    double C[9];
    C[0]     = F[4] * F[8] - F[5] * F[7];
    C[1]     = F[5] * F[6] - F[3] * F[8];
    C[2]     = F[3] * F[7] - F[4] * F[6];
    C[3]     = F[2] * F[7] - F[1] * F[8];
    C[4]     = F[0] * F[8] - F[2] * F[6];
    C[5]     = F[1] * F[6] - F[0] * F[7];
    C[6]     = F[1] * F[5] - F[2] * F[4];
    C[7]     = F[2] * F[3] - F[0] * F[5];
    C[8]     = F[0] * F[4] - F[1] * F[3];
    double J = F[0] * C[0] + F[1] * C[1] + F[2] * C[2];
    double lnJ = log(J);
    double cP  = (lam * lnJ - mu) / J;
    double P[9];
#pragma unroll
    for(int k = 0; k < 9; k++)
        P[k] = mu * F[k] + cP * C[k];
    // B[a][j] = dF_ij / dx_{a,i}
    double B[12];
#pragma unroll
    for(int j = 0; j < 3; j++)
    {
        B[j]         = -(Dm[j] + Dm[3 + j] + Dm[6 + j]);
        B[3 + j]     = Dm[j];
        B[6 + j]     = Dm[3 + j];
        B[9 + j]     = Dm[6 + j];
    }
    double  s  = dt2 * vol[e];
    double* ob = out + (size_t)kFemStride * e;
#pragma unroll
    for(int a = 0; a < 4; a++)
#pragma unroll
        for(int i = 0; i < 3; i++)
            ob[90 + 3 * a + i] = s * (P[i * 3 + 0] * B[a * 3 + 0] + P[i * 3 + 1] * B[a * 3 + 1]
                                      + P[i * 3 + 2] * B[a * 3 + 2]);
    if(!hessian)
        return;

    // ---- analytic eigensystem of the SNK1 Hessian in F space
    double U[9], S[3], Vm[9];
    svd3(F, U, S, Vm);
    double H[78];
#pragma unroll
    for(int k = 0; k < 78; k++)
        H[k] = 0.0;
    const double isq2 = 0.70710678118654752440;
    auto rank1 = [&](const double Q[9], double lam_k) {
        if(!(lam_k > 0.0))
            return;
        double w[12];
#pragma unroll
        for(int a = 0; a < 4; a++)
#pragma unroll
            for(int i = 0; i < 3; i++)
                w[3 * a + i] = Q[i * 3 + 0] * B[a * 3 + 0] + Q[i * 3 + 1] * B[a * 3 + 1]
                               + Q[i * 3 + 2] * B[a * 3 + 2];
        double sl = s * lam_k;
#pragma unroll
        for(int p = 0; p < 12; p++)
        {
            double wp = sl * w[p];
#pragma unroll
            for(int q = p; q < 12; q++)
                H[uidx(p, q)] += wp * w[q];
        }
    };
    // twist / flip modes: (u_i v_j^T -/+ u_j v_i^T) / sqrt2, eigenvalue mu +/- cP sigma_k
#pragma unroll
    for(int k = 0; k < 3; k++)
    {
        int    i = (k == 0) ? 1 : 0;
        int    j = (k == 2) ? 1 : 2;
        double Qt[9], Qf[9];
#pragma unroll
        for(int r = 0; r < 3; r++)
#pragma unroll
            for(int c = 0; c < 3; c++)
            {
                double a = U[r * 3 + i] * Vm[c * 3 + j], b = U[r * 3 + j] * Vm[c * 3 + i];
                Qt[r * 3 + c] = (a - b) * isq2;
                Qf[r * 3 + c] = (a + b) * isq2;
            }
        rank1(Qt, mu + cP * S[k]);
        rank1(Qf, mu - cP * S[k]);
    }
    // scaling modes: U diag(a) V^T, a eigenvectors of the 3x3 block
    {
        double A3[9], w3[3], E[9];
        double tmp = lam * (1 - lnJ) + mu;
        A3[0]      = mu + tmp / (S[0] * S[0]);
        A3[4]      = mu + tmp / (S[1] * S[1]);
        A3[8]      = mu + tmp / (S[2] * S[2]);
        A3[1] = A3[3] = lam / (S[0] * S[1]);
        A3[2] = A3[6] = lam / (S[0] * S[2]);
        A3[5] = A3[7] = lam / (S[1] * S[2]);
        sym_eig3(A3, w3, E);
#pragma unroll
        for(int k = 0; k < 3; k++)
        {
            double Q[9];
#pragma unroll
            for(int r = 0; r < 3; r++)
#pragma unroll
                for(int c = 0; c < 3; c++)
                    Q[r * 3 + c] = U[r * 3 + 0] * E[0 * 3 + k] * Vm[c * 3 + 0]
                                   + U[r * 3 + 1] * E[1 * 3 + k] * Vm[c * 3 + 1]
                                   + U[r * 3 + 2] * E[2 * 3 + k] * Vm[c * 3 + 2];
            rank1(Q, w3[k]);
        }
    }
    // ---- the 10 upper blocks (a <= b), row-major
#pragma unroll
    for(int pi = 0; pi < 10; pi++)
    {
        const int a = kPairA[pi], b = kPairB[pi];
#pragma unroll
        for(int r = 0; r < 3; r++)
#pragma unroll
            for(int c = 0; c < 3; c++)
            {
                int p = 3 * a + r, q = 3 * b + c;
                ob[9 * pi + 3 * r + c] = (p <= q) ? H[uidx(p, q)] : H[uidx(q, p)];
            }
    }
}

// One thread per upper BSR block (i <= j): sum of the element blocks listed
// for it (code = (buffer offset / 3) << 1 | transposed) added into (i, j)
// and, for i != j, its transpose into (j, i).
__global__ void __launch_bounds__(256) k_fem_gather_hess(int n_up, const int* __restrict__ up,
                                                         const int* __restrict__ lo,
                                                         const int* __restrict__ ptr,
                                                         const int* __restrict__ src,
                                                         const double* __restrict__ buf,
                                                         double* __restrict__ vals)
{
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if(u >= n_up)
        return;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for(int k = ptr[u]; k < ptr[u + 1]; k++)
    {
        int           code = src[k];
        const double* b    = buf + 3 * (size_t)(code >> 1);
        if(code & 1)
        {
#pragma unroll
            for(int r = 0; r < 3; r++)
#pragma unroll
                for(int c = 0; c < 3; c++)
                    acc[3 * r + c] += b[3 * c + r];
        }
        else
        {
#pragma unroll
            for(int q = 0; q < 9; q++)
                acc[q] += b[q];
        }
    }
    double* d = vals + 9 * (size_t)up[u];
#pragma unroll
    for(int q = 0; q < 9; q++)
        d[q] += acc[q];
    int l = lo[u];
    if(l != up[u])
    {
        double* dt = vals + 9 * (size_t)l;
#pragma unroll
        for(int r = 0; r < 3; r++)
#pragma unroll
            for(int c = 0; c < 3; c++)
                dt[3 * r + c] += acc[3 * c + r];
    }
}

// One thread per vertex: the element gradient entries listed for it
// (code = buffer offset / 3, those below code_lim) added into the shape
// gradient.
__global__ void __launch_bounds__(256) k_fem_gather_grad(int nv, const int* __restrict__ ptr,
                                                         const int* __restrict__ src,
                                                         const double* __restrict__ buf,
                                                         double* __restrict__ gshape,
                                                         int code_lim)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if(v >= nv)
        return;
    double g0 = 0, g1 = 0, g2 = 0;
    for(int k = ptr[v]; k < ptr[v + 1]; k++)
    {
        int           code = src[k];
        if(code >= code_lim)
            continue;
        const double* b    = buf + 3 * (size_t)code;
        g0 += b[0];
        g1 += b[1];
        g2 += b[2];
    }
    gshape[3 * (size_t)v] += g0;
    gshape[3 * (size_t)v + 1] += g1;
    gshape[3 * (size_t)v + 2] += g2;
}

// ----------------------------------------------------------------- ground
__global__ void k_ground_detect(int ns, const uint32_t* __restrict__ surf,
                                const double* __restrict__ x, V3 n, double off, double dHat,
                                uint32_t* __restrict__ gp, unsigned* __restrict__ gcount)
{
    int  i    = blockIdx.x * blockDim.x + threadIdx.x;
    bool pred = false;
    uint32_t v = 0;
    if(i < ns)
    {
        v           = surf[i];
        double dist = dot(n, ld3(x + 3 * v)) - off;
        pred        = !(dist * dist > dHat);
    }
    unsigned slot = wave_append(gcount, pred);
    if(pred)
        gp[slot] = v;
}

__global__ void k_ground(const unsigned* __restrict__ gcount, const uint32_t* __restrict__ gp,
                         const double* __restrict__ x, V3 n, double off, double dHat,
                         double kappa, const int* __restrict__ btype, BlockSink sink,
                         double* __restrict__ gcontact, int hessian)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= *gcount)
        return;
    uint32_t v     = gp[i];
    double   dist  = dot(n, ld3(x + 3 * v)) - off;
    double   dist2 = dist * dist;
    double   t     = dist2 - dHat;
    double   g_b   = t * log(dist2 / dHat) * -2.0 - (t * t) / dist2;
    double   H_b   = (log(dist2 / dHat) * -2.0 - t * 4.0 / dist2) + 1.0 / (dist2 * dist2) * (t * t);
    double   gs    = kappa * g_b * 2 * dist;
    atomicAdd(gcontact + 3 * v + 0, n.x * gs);
    atomicAdd(gcontact + 3 * v + 1, n.y * gs);
    atomicAdd(gcontact + 3 * v + 2, n.z * gs);
    if(!hessian || btype[v] != 0)
        return;
    double p      = kappa * (4.0 * H_b * dist2 + 2.0 * g_b);
    double nn[3]  = {n.x, n.y, n.z};
    double blk[9];
    for(int r = 0; r < 3; r++)
        for(int c = 0; c < 3; c++)
            blk[r * 3 + c] = p * nn[r] * nn[c];
    sink.put((int)v, (int)v, blk, true);
}

// --------------------------------------------------------------- barrier
__device__ __forceinline__ double pair_d2(int type, const double* x, const int ids[4])
{
    V3 a = ld3(x + 3 * ids[0]), b = ld3(x + 3 * ids[1]);
    if(type == 0)
        return d_pp(a, b);
    V3 c = ld3(x + 3 * ids[2]);
    if(type == 1)
        return d_pe(a, b, c);
    V3 d = ld3(x + 3 * ids[3]);
    return type == 2 ? d_pt(a, b, c, d) : d_ee(a, b, c, d);
}

// gradient of the unsigned distance (closest points; envelope theorem):
// g_k = w_k u for the element's vertices k < nv, u the unit direction
// between the closest points; returns nv
__device__ inline int dist_wu(int type, const double* x, const int ids[4], double w[4], V3& u)
{
    V3 x0 = ld3(x + 3 * ids[0]), x1 = ld3(x + 3 * ids[1]);
    if(type == 0)
    {
        u    = x0 - x1;
        u    = u * (1.0 / sqrt(sqn(u)));
        w[0] = 1.0, w[1] = -1.0;
        return 2;
    }
    V3 x2 = ld3(x + 3 * ids[2]);
    if(type == 1)
    {
        V3     e = x2 - x1;
        double t = dot(x0 - x1, e) / sqn(e);
        u        = x0 - (x1 + e * t);
        u        = u * (1.0 / sqrt(sqn(u)));
        w[0] = 1.0, w[1] = -(1.0 - t), w[2] = -t;
        return 3;
    }
    V3 x3 = ld3(x + 3 * ids[3]);
    if(type == 2)
    {
        V3     e1 = x2 - x1, e2 = x3 - x1, q = x0 - x1;
        V3     nv = cross(e1, e2);
        double h  = dot(q, nv) / sqn(nv);
        V3     pq = q - nv * h;
        double a = sqn(e1), b = dot(e1, e2), c = sqn(e2);
        double r1 = dot(pq, e1), r2 = dot(pq, e2), D = a * c - b * b;
        double be = (c * r1 - b * r2) / D, ga = (a * r2 - b * r1) / D;
        u         = nv * h;
        u         = u * (1.0 / sqrt(sqn(u)));
        w[0] = 1.0, w[1] = -(1.0 - be - ga), w[2] = -be, w[3] = -ga;
        return 4;
    }
    V3     eu = x1 - x0, v = x3 - x2, ww = x0 - x2;
    double a = sqn(eu), b = dot(eu, v), c = sqn(v), d = dot(eu, ww), e = dot(v, ww);
    double D = a * c - b * b;
    double s = (b * e - c * d) / D, t = (a * e - b * d) / D;
    u        = ww + eu * s - v * t;
    u        = u * (1.0 / sqrt(sqn(u)));
    w[0] = 1.0 - s, w[1] = s, w[2] = -(1.0 - t), w[3] = -t;
    return 4;
}

__device__ __forceinline__ double barrier_lambda0(double I5, double dHat, double kappa)
{
    double L = log(I5);
    return -(4 * kappa * dHat * dHat
             * (4 * I5 + L - 3 * I5 * I5 * L * L + 6 * I5 * L - 2 * I5 * I5 + I5 * L * L
                - 7 * I5 * I5 * L - 2))
           / I5;
}

__global__ void __launch_bounds__(128) k_barrier(const unsigned* __restrict__ count, int pair_cap,
                                                 const int4* __restrict__ pairs,
                                                 const double* __restrict__ x, double dHat,
                                                 double kappa, BlockSink sink,
                                                 double* __restrict__ gcontact, int hessian)
{
    unsigned pi = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned np = *count;
    if(pi >= (np < (unsigned)pair_cap ? np : (unsigned)pair_cap))
        return;
    int    ids[4];
    int    type = decode_pair(pairs[pi], ids);
    int    nv   = (type == 0) ? 2 : (type == 1) ? 3 : 4;
    double dis  = pair_d2(type, x, ids);
    double d    = sqrt(dis);
    double dhs  = sqrt(dHat);
    double sq   = d / dhs;
    double I5   = sq * sq;
    double L    = log(I5);
    double pk1  = sq * 2 * (2 * kappa * dHat * dHat * L * (I5 - 1) * (I5 + I5 * L - 1)) / I5;
    // d(sqrt(I5))/dx = g / sqrt(dhat), g_k = w_k u
    double w[4];
    V3     u;
    dist_wu(type, x, ids, w, u);
    double cw[4];
    for(int a = 0; a < nv; a++)
        cw[a] = w[a] * (1.0 / dhs);
    const double uu[3] = {u.x, u.y, u.z};
    for(int a = 0; a < nv; a++)
        for(int c = 0; c < 3; c++)
            atomicAdd(gcontact + 3 * ids[a] + c, pk1 * (cw[a] * uu[c]));
    if(!hessian)
        return;
    double lam0 = barrier_lambda0(I5, dHat, kappa);
    if(d * d < 1e-6 * dHat)
        lam0 = barrier_lambda0(1e-6, dHat, kappa);
    // H_ab = lambda0 (dsq/dx_a)(dsq/dx_b)^T = cw_a cw_b (lambda0 u u^T)
    double K[9];
    for(int r = 0; r < 3; r++)
        for(int c = 0; c < 3; c++)
            K[r * 3 + c] = lam0 * uu[r] * uu[c];
    sink.put_rank(ids, nv, cw, K);
}

__global__ void k_rhs(int nv, const int* __restrict__ btype, const double* __restrict__ gc,
                      const double* __restrict__ gs, double* __restrict__ rhs)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= 3 * nv)
        return;
    rhs[i] = btype[i / 3] != 0 ? 0.0 : gc[i] + gs[i];
}

// --------------------------------------------------------------- energies
__global__ void __launch_bounds__(256) k_energy_kin(int nv, const double* __restrict__ x,
                                                    const double* __restrict__ xt,
                                                    const double* __restrict__ mass, double* out)
{
    int    i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if(i < nv)
    {
        V3 d = ld3(x + 3 * i) - ld3(xt + 3 * i);
        e    = sqn(d) * mass[i] * 0.5;
    }
    block_partial(e, out);
}

__global__ void __launch_bounds__(256) k_energy_fem(int nt, const uint32_t* __restrict__ tets,
                                                    const double* __restrict__ dminv,
                                                    const double* __restrict__ vol,
                                                    const double* __restrict__ mu_,
                                                    const double* __restrict__ lam_,
                                                    const double* __restrict__ x, double dt2,
                                                    double* out)
{
    int    e = blockIdx.x * blockDim.x + threadIdx.x;
    double en = 0.0;
    if(e < nt)
    {
        const uint32_t* t  = tets + 4 * (size_t)e;
        V3              x0 = ld3(x + 3 * t[0]);
        V3 d1 = ld3(x + 3 * t[1]) - x0, d2 = ld3(x + 3 * t[2]) - x0, d3 = ld3(x + 3 * t[3]) - x0;
        double Ds[9] = {d1.x, d2.x, d3.x, d1.y, d2.y, d3.y, d1.z, d2.z, d3.z};
        double Dm[9], F[9];
        for(int k = 0; k < 9; k++)
            Dm[k] = dminv[(size_t)k * nt + e];
        for(int i = 0; i < 3; i++)
            for(int j = 0; j < 3; j++)
                F[i * 3 + j] = Ds[i * 3 + 0] * Dm[j] + Ds[i * 3 + 1] * Dm[3 + j]
                               + Ds[i * 3 + 2] * Dm[6 + j];
        double I1 = 0;
        for(int k = 0; k < 9; k++)
            I1 += F[k] * F[k];
        double J = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6])
                   + F[2] * (F[3] * F[7] - F[4] * F[6]);
        double lj = log(J);
        double mu = mu_[e], lam = lam_[e];
        en = dt2 * (0.5 * mu * (I1 - 3) - mu * lj + 0.5 * lam * lj * lj) * vol[e];
    }
    block_partial(en, out);
}

__global__ void __launch_bounds__(256) k_energy_barrier(const unsigned* __restrict__ count,
                                                        int pair_cap,
                                                        const int4* __restrict__ pairs,
                                                        const double* __restrict__ x,
                                                        double dHat, double kappa, double* out)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    double   e = 0.0;
    unsigned np = *count;
    if(i < (np < (unsigned)pair_cap ? np : (unsigned)pair_cap))
    {
        int    ids[4];
        int    type = decode_pair(pairs[i], ids);
        double dis  = pair_d2(type, x, ids);
        double I5   = dis / dHat;
        double lenE = dis - dHat;
        e           = kappa * lenE * lenE * log(I5) * log(I5);
    }
    block_partial(e, out);
}

__global__ void __launch_bounds__(256) k_energy_ground(const unsigned* __restrict__ gcount,
                                                       const uint32_t* __restrict__ gp,
                                                       const double* __restrict__ x, V3 n,
                                                       double off, double dHat, double kappa,
                                                       double* out)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    double   e = 0.0;
    if(i < *gcount)
    {
        double dist  = dot(n, ld3(x + 3 * gp[i])) - off;
        double dist2 = dist * dist;
        e            = kappa * (-(dist2 - dHat) * (dist2 - dHat) * log(dist2 / dHat));
    }
    block_partial(e, out);
}

// _computeSoftConstraintEnergy_Reduction (GIPC.cu:7527-7560): d |rate (x - t)|^2 / 2
__global__ void __launch_bounds__(256) k_energy_soft(int n, const uint32_t* __restrict__ idx,
                                                     const double* __restrict__ tgt,
                                                     const double* __restrict__ x, double d,
                                                     double rate, double* out)
{
    int    i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if(i < n)
    {
        V3 dv = (ld3(x + 3 * (size_t)idx[i]) - ld3(tgt + 3 * (size_t)i)) * rate;
        e     = d * sqn(dv) * 0.5;
    }
    block_partial(e, out);
}

// sum the partial arrays (fixed order) into terms[0..9]; terms[10] = total
__global__ void __launch_bounds__(256) k_energy_final(const double* __restrict__ partial,
                                                      EnergyLayout lay, double* __restrict__ terms)
{
    __shared__ double ws[4];
    for(int term = 0; term < kEnergyTerms; term++)
    {
        double v = 0.0;
        for(int i = threadIdx.x; i < lay.n[term]; i += blockDim.x)
            v += partial[lay.off[term] + i];
        v = wave_sum(v);
        if(__lane_id() == 0)
            ws[threadIdx.x >> 6] = v;
        __syncthreads();
        if(threadIdx.x == 0)
            terms[term] = ws[0] + ws[1] + ws[2] + ws[3];
        __syncthreads();
    }
    if(threadIdx.x == 0)
    {
        double t = 0.0;
        for(int term = 0; term < kEnergyTerms; term++)
            t += terms[term];
        terms[kEnergyTerms] = t;
    }
}

// ---------------------------------------------------------------- helpers
__global__ void k_dot3(int n, const double* __restrict__ a, const double* __restrict__ b,
                       double* out)
{
    int    i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = (i < n) ? a[i] * b[i] : 0.0;
    d        = wave_sum(d);
    if(__lane_id() == 0 && d != 0.0)
        atomicAdd(out, d);
}

// grid-stride max |a_i|: one atomic per workgroup (a single address would
// serialise one atomic per wave)
__global__ void __launch_bounds__(256) k_max_abs(int n, const double* __restrict__ a, double* out)
{
    __shared__ double wm[4];
    double            d = 0.0;
    for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        d = fmax(d, fabs(a[i]));
    d = wave_max(d);
    if(__lane_id() == 0)
        wm[threadIdx.x >> 6] = d;
    __syncthreads();
    if(threadIdx.x == 0)
    {
        d = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
        if(d > 0.0)
            atomic_max_pos(out, d);
    }
}

__global__ void k_step_forward(int nv, const int* __restrict__ btype, const double* __restrict__ xs,
                               const double* __restrict__ dx, double alpha, double* __restrict__ x)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= 3 * nv)
        return;
    if(btype[i / 3] == 0)
        x[i] = xs[i] - dx[i] * alpha;
}

__global__ void k_update_velocity(int nv, const int* __restrict__ btype, const double* __restrict__ x,
                                  double* __restrict__ xprev, double* __restrict__ v, double dt,
                                  V3 g, double* __restrict__ xt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= nv)
        return;
    bool   fr = btype[i] == 0;
    double gg[3] = {g.x, g.y, g.z};
    for(int c = 0; c < 3; c++)
    {
        size_t k  = 3 * (size_t)i + c;
        double vv = fr ? (x[k] - xprev[k]) * (1 / dt) : 0.0;
        v[k]      = vv;
        xprev[k]  = x[k];
        xt[k]     = xprev[k] + (vv * dt + (fr ? gg[c] * (dt * dt) : 0.0));
    }
}

// ----------------------------------------------------------------- launch
void launch_init_diag(const SimDev& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_init_diag, dim3(grid_for(d.nv, 256)), dim3(256), 0, s, d.nv, d.nfem,
                       d.mass, d.x, d.xt, d.diag, d.vals, d.gshape, d.gcontact);
    SG_LAUNCH_CHECK();
}

void launch_soft(const SimDev& d, bool hessian, double* gshape, hipStream_t s)
{
    if(d.nsoft == 0)
        return;
    hipLaunchKernelGGL(k_soft, dim3(grid_for(d.nsoft, 256)), dim3(256), 0, s, d.nsoft,
                       d.soft_idx, d.soft_tgt, d.x, d.soft_k, d.soft_rate, d.diag, d.vals,
                       gshape, hessian ? 1 : 0);
    SG_LAUNCH_CHECK();
}

void launch_fem(const SimDev& d, bool hessian, hipStream_t s, bool cloth)
{
    if(d.nt == 0 && (!cloth || (d.cloth.nf == 0 && d.cloth.nh == 0)))
        return;
    if(d.nt)
    {
        hipLaunchKernelGGL(k_fem, dim3(grid_for(d.nt, 128)), dim3(128), 0, s, d.nt, d.tets,
                           d.dminv, d.vol, d.mu, d.lam, d.x, d.dt * d.dt, d.fem_buf,
                           hessian ? 1 : 0);
        SG_LAUNCH_CHECK();
    }
    if(cloth)
        launch_cloth(d, hessian, s);
    hipLaunchKernelGGL(k_fem_gather_grad, dim3(grid_for(d.nv, 256)), dim3(256), 0, s, d.nv,
                       d.gv_ptr, d.gv_src, d.fem_buf, d.gshape,
                       cloth ? INT_MAX : (int)(d.cloth.off_tri / 3));
    SG_LAUNCH_CHECK();
    if(hessian && d.n_up > 0)
    {
        hipLaunchKernelGGL(k_fem_gather_hess, dim3(grid_for(d.n_up, 256)), dim3(256), 0, s,
                           d.n_up, d.gh_up, d.gh_lo, d.gh_ptr, d.gh_src, d.fem_buf, d.vals);
        SG_LAUNCH_CHECK();
    }
}

void launch_ground_detect(const SimDev& d, hipStream_t s)
{
    SG_HIP(hipMemsetAsync(d.gcount, 0, sizeof(unsigned), s));
    if(!d.use_ground || d.ns == 0)
        return;
    hipLaunchKernelGGL(k_ground_detect, dim3(grid_for(d.ns, 256)), dim3(256), 0, s, d.ns, d.surf,
                       d.x, d.gn, d.goff, d.dHat, d.gp, d.gcount);
    SG_LAUNCH_CHECK();
}

void launch_ground(const SimDev& d, double kappa, bool hessian, double* gcontact, hipStream_t s)
{
    if(!d.use_ground || d.ns == 0)
        return;
    hipLaunchKernelGGL(k_ground, dim3(grid_for(d.ns, 256)), dim3(256), 0, s, d.gcount, d.gp, d.x,
                       d.gn, d.goff, d.dHat, kappa, d.btype, block_sink(d), gcontact,
                       hessian ? 1 : 0);
    SG_LAUNCH_CHECK();
}

void launch_barrier(const SimDev& d, double kappa, bool hessian, double* gcontact, hipStream_t s)
{
    if(d.pair_cap == 0)
        return;
    hipLaunchKernelGGL(k_barrier, dim3(grid_for(d.pair_cap, 128)), dim3(128), 0, s, d.pcount,
                       d.pair_cap, d.pairs, d.x, d.dHat, kappa, block_sink(d), gcontact,
                       hessian ? 1 : 0);
    SG_LAUNCH_CHECK();
}

void launch_rhs(const SimDev& d, double* rhs, hipStream_t s)
{
    if(d.nfem == 0)
        return;
    hipLaunchKernelGGL(k_rhs, dim3(grid_for(3 * (long long)d.nfem, 256)), dim3(256), 0, s, d.nfem,
                       d.btype, d.gcontact, d.gshape, rhs);
    SG_LAUNCH_CHECK();
}

EnergyLayout energy_layout(const SimDev& d)
{
    EnergyLayout L;
    L.n[0]   = (int)grid_for(d.nv, 256);
    L.n[1]   = d.nt ? (int)grid_for(d.nt, 256) : 0;
    L.n[2]   = d.pair_cap ? (int)grid_for(d.pair_cap, 256) : 0;
    L.n[3]   = (d.use_ground && d.ns) ? (int)grid_for(d.ns, 256) : 0;
    L.n[4]   = (d.fric.mu != 0.0 && d.fric.cap) ? (int)grid_for(d.fric.cap, 256) : 0;
    L.n[5]   = (d.fric.mu_gd != 0.0 && d.use_ground && d.ns) ? (int)grid_for(d.ns, 256) : 0;
    L.n[6]   = d.cloth.nf ? (int)grid_for(d.cloth.nf, 256) : 0;
    L.n[7]   = d.cloth.nh ? (int)grid_for(d.cloth.nh, 256) : 0;
    L.n[8]   = d.abd.nb ? (int)grid_for(d.abd.nb, 256) : 0;
    L.n[9]   = d.nsoft ? (int)grid_for(d.nsoft, 256) : 0;
    L.off[0] = 0;
    for(int k = 1; k < kEnergyTerms; k++)
        L.off[k] = L.off[k - 1] + L.n[k - 1];
    return L;
}

void launch_energy(const SimDev& d, double kappa, double* partial, double* terms, hipStream_t s)
{
    EnergyLayout L = energy_layout(d);
    hipLaunchKernelGGL(k_energy_kin, dim3(L.n[0]), dim3(256), 0, s, d.nv, d.x, d.xt, d.mass,
                       partial + L.off[0]);
    if(L.n[1])
        hipLaunchKernelGGL(k_energy_fem, dim3(L.n[1]), dim3(256), 0, s, d.nt, d.tets, d.dminv,
                           d.vol, d.mu, d.lam, d.x, d.dt * d.dt, partial + L.off[1]);
    if(L.n[2])
        hipLaunchKernelGGL(k_energy_barrier, dim3(L.n[2]), dim3(256), 0, s, d.pcount,
                           d.pair_cap, d.pairs,
                           d.x, d.dHat, kappa, partial + L.off[2]);
    if(L.n[3])
        hipLaunchKernelGGL(k_energy_ground, dim3(L.n[3]), dim3(256), 0, s, d.gcount, d.gp, d.x,
                           d.gn, d.goff, d.dHat, kappa, partial + L.off[3]);
    launch_energy_friction(d, L, partial, s);
    launch_energy_cloth(d, L, partial, s);
    launch_energy_abd(d, L, partial, s);
    if(L.n[9])
        hipLaunchKernelGGL(k_energy_soft, dim3(L.n[9]), dim3(256), 0, s, d.nsoft, d.soft_idx,
                           d.soft_tgt, d.x, d.soft_k, d.soft_rate, partial + L.off[9]);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, s, partial, L, terms);
    SG_LAUNCH_CHECK();
}

void launch_dot(int n, const double* a, const double* b, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_dot3, dim3(grid_for(n, 256)), dim3(256), 0, s, n, a, b, out);
    SG_LAUNCH_CHECK();
}

void launch_max_abs(int n, const double* a, double* out, hipStream_t s)
{
    unsigned g = grid_for(n, 256);
    hipLaunchKernelGGL(k_max_abs, dim3(g < 1024 ? g : 1024), dim3(256), 0, s, n, a, out);
    SG_LAUNCH_CHECK();
}

void launch_step_forward(const SimDev& d, const double* xs, const double* dx, double alpha,
                         hipStream_t s)
{
    if(d.nfem)
        hipLaunchKernelGGL(k_step_forward, dim3(grid_for(3 * (long long)d.nfem, 256)), dim3(256),
                           0, s, d.nfem, d.btype, xs, dx, alpha, d.x);
    SG_LAUNCH_CHECK();
}

void launch_update_velocity(const SimDev& d, double* xprev, double* v, hipStream_t s)
{
    hipLaunchKernelGGL(k_update_velocity, dim3(grid_for(d.nv, 256)), dim3(256), 0, s, d.nv,
                       d.btype, d.x, xprev, v, d.dt, d.grav, d.xt);
    SG_LAUNCH_CHECK();
}

}  // namespace sgipc
