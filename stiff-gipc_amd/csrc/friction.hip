// friction.hip -- lagged IPC friction: self contacts and ground.
//
//   k_friction_build        lagged normal force, closest-point coordinates and
//                           tangent basis per contact at the start of a time
//                           step (_calFrictionLastH_DistAndTan GIPC.cu:8330-8441,
//                           buildFrictionSets :8601-8637; FrictionUtils.cuh)
//   k_friction_ground_build ground lagged force (_calFrictionLastH_gd :8304-8328)
//   k_friction              gradient mu lam T f1(|u|)/|u| u and Hessian
//                           mu lam T PSD(M2) T^T (_calFrictionGradient :5089-5273,
//                           _calFrictionHessian :1441-1804)
//   k_friction_ground       ground gradient / Hessian (:5049-5088, :1359-1440)
//   k_energy_friction*      energies (:1254-1352), C1 static mollifier
//                           (SFCLAMPING_ORDER 1, FrictionUtils.cuh:419-499)
//
// u is the tangential relative displacement since the step start, projected
// on the lagged basis; eps^2 = 1e-4 bboxDiag^2 dt^2 (GIPC.cu:8544, :9128).
// Only non-mollified contacts carry friction (EE with w >= 0, PP/PE with
// y >= 0, PT).  The ground Hessian keeps the reference's x/z tangent plane
// (it assumes the ground normal (0, 1, 0)).
#include "assembly.h"

namespace sgipc
{

// ------------------------------------------------------------ geometry
__device__ __forceinline__ V3 nrmz(V3 a) { return a * (1.0 / sqrt(sqn(a))); }

// basis columns b0, b1 stored row-major 3x2: B[2 r + c]
__device__ __forceinline__ void put_basis(double* B, V3 c0, V3 c1)
{
    B[0] = c0.x, B[2] = c0.y, B[4] = c0.z;
    B[1] = c1.x, B[3] = c1.y, B[5] = c1.z;
}

__device__ __forceinline__ void inv2(const double m[4], double o[4])
{
    double det = m[0] * m[3] - m[1] * m[2];
    o[0]       = m[3] / det;
    o[1]       = -m[1] / det;
    o[2]       = -m[2] / det;
    o[3]       = m[0] / det;
}

// kind 0 PP, 1 PE, 2 PT, 3 EE, -1 no friction
__device__ __forceinline__ int friction_kind(int4 c, int ids[4])
{
    if(c.x >= 0)
    {
        ids[0] = c.x, ids[1] = c.y, ids[2] = c.z, ids[3] = c.w;
        return c.w < 0 ? -1 : 3;
    }
    ids[0] = -c.x - 1;
    ids[1] = c.y;
    ids[2] = c.z;
    ids[3] = c.w;
    if(c.z < 0)
        return c.y >= 0 ? 0 : -1;
    if(c.w < 0)
        return c.y >= 0 ? 1 : -1;
    return 2;
}

__device__ __forceinline__ int kind_weights(int kind, const double c[2], double w[4])
{
    if(kind == 0)
    {
        w[0] = 1.0, w[1] = -1.0;
        return 2;
    }
    if(kind == 1)
    {
        w[0] = 1.0, w[1] = c[0] - 1.0, w[2] = -c[0];
        return 3;
    }
    if(kind == 2)
    {
        w[0] = 1.0, w[1] = -1.0 + c[0] + c[1], w[2] = -c[0], w[3] = -c[1];
        return 4;
    }
    w[0] = 1.0 - c[0], w[1] = c[0], w[2] = c[1] - 1.0, w[3] = -c[1];
    return 4;
}

// lagged force, coordinates and basis; returns the kind
__device__ int friction_lag(const double* x, int4 code, double dHat, double kappa, double* lam,
                            double c[2], double B[6])
{
    int ids[4];
    int kind = friction_kind(code, ids);
    if(kind < 0)
        return -1;
    V3     v0 = ld3(x + 3 * ids[0]), v1 = ld3(x + 3 * ids[1]);
    double dis;
    c[0] = c[1] = 0.0;
    if(kind == 0)
    {
        dis    = d_pp(v0, v1);
        V3 v01 = v1 - v0;
        V3 xc  = V3{0.0, -v01.z, v01.y};
        V3 yc  = V3{v01.z, 0.0, -v01.x};
        if(sqn(xc) > sqn(yc))
            put_basis(B, nrmz(xc), nrmz(cross(v01, xc)));
        else
            put_basis(B, nrmz(yc), nrmz(cross(v01, yc)));
    }
    else if(kind == 1)
    {
        V3 v2  = ld3(x + 3 * ids[2]);
        dis    = d_pe(v0, v1, v2);
        V3 e12 = v2 - v1;
        c[0]   = dot(v0 - v1, e12) / sqn(e12);
        put_basis(B, nrmz(e12), nrmz(cross(e12, v0 - v1)));
    }
    else
    {
        V3 v2 = ld3(x + 3 * ids[2]), v3 = ld3(x + 3 * ids[3]);
        if(kind == 2)
        {
            dis     = d_pt(v0, v1, v2, v3);
            V3 v12 = v2 - v1, v13 = v3 - v1, v10 = v0 - v1;
            double m[4] = {sqn(v12), dot(v12, v13), dot(v13, v12), sqn(v13)}, mi[4];
            inv2(m, mi);
            double b0 = dot(v12, v10), b1 = dot(v13, v10);
            c[0]      = mi[0] * b0 + mi[1] * b1;
            c[1]      = mi[2] * b0 + mi[3] * b1;
            put_basis(B, nrmz(v12), nrmz(cross(cross(v12, v13), v12)));
        }
        else
        {
            dis    = d_ee(v0, v1, v2, v3);
            V3 e20 = v0 - v2, e01 = v1 - v0, e23 = v3 - v2;
            double m[4] = {sqn(e01), -dot(e23, e01), -dot(e23, e01), sqn(e23)}, mi[4];
            inv2(m, mi);
            double r0 = -dot(e20, e01), r1 = dot(e20, e23);
            c[0]      = mi[0] * r0 + mi[1] * r1;
            c[1]      = mi[2] * r0 + mi[3] * r1;
            put_basis(B, nrmz(e01), nrmz(cross(cross(e01, e23), e01)));
        }
    }
    double L = log(dis / dHat);
    *lam     = -kappa * 2.0 * sqrt(dis)
           * (L * L * (2 * dis - 2 * dHat) + (2 * L * (dis - dHat) * (dis - dHat)) / dis);
    return kind;
}

// u = B^T sum_k w_k (x_k - xprev_k)
__device__ __forceinline__ int rel_u(const double* x, const double* xp, int4 code,
                                     const double c[2], const double B[6], int ids[4],
                                     double w[4], double u[2])
{
    int kind = friction_kind(code, ids);
    int m    = kind_weights(kind, c, w);
    V3  r    = V3{0.0, 0.0, 0.0};
    for(int k = 0; k < m; k++)
        r = r + (ld3(x + 3 * ids[k]) - ld3(xp + 3 * ids[k])) * w[k];
    u[0] = B[0] * r.x + B[2] * r.y + B[4] * r.z;
    u[1] = B[1] * r.x + B[3] * r.y + B[5] * r.z;
    return m;
}

__device__ __forceinline__ double f0_c1(double x2, double eps)
{
    return x2 * (-sqrt(x2) / 3.0 + eps) / (eps * eps) + eps / 3.0;
}
__device__ __forceinline__ double f1_div_c1(double x2, double eps)
{
    return (-sqrt(x2) + 2.0 * eps) / (eps * eps);
}
__device__ __forceinline__ double f2_c1(double x2, double eps)
{
    return 2.0 * (eps - sqrt(x2)) / (eps * eps);
}

// symmetric 2x2 PSD projection, unchanged when both eigenvalues are >= 0
__device__ __forceinline__ void psd2(double M[4])
{
    double a = M[0], b = 0.5 * (M[1] + M[2]), d = M[3];
    double tr = 0.5 * (a + d), df = 0.5 * (a - d);
    double r  = sqrt(df * df + b * b);
    if(tr - r >= 0.0)
        return;
    double l1 = tr + r, vx, vy;
    if(r == 0.0)
        vx = 1.0, vy = 0.0;
    else if(df >= 0.0)
        vx = df + r, vy = b;
    else
        vx = b, vy = r - df;
    double nv = sqrt(vx * vx + vy * vy);
    vx /= nv, vy /= nv;
    double k = l1 > 0.0 ? l1 : 0.0;
    M[0]     = k * vx * vx;
    M[1] = M[2] = k * vx * vy;
    M[3]        = k * vy * vy;
}

// --------------------------------------------------------------- kernels
__global__ void __launch_bounds__(256) k_friction_build(const unsigned* __restrict__ pcount,
                                                        int pair_cap,
                                                        const int4* __restrict__ pairs,
                                                        const double* __restrict__ x,
                                                        double dHat, double kappa, FricDev f)
{
    unsigned i  = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned np = *pcount;
    np          = np < (unsigned)pair_cap ? np : (unsigned)pair_cap;
    double lam = 0.0, c[2] = {0.0, 0.0}, B[6] = {0, 0, 0, 0, 0, 0};
    int4   code = make_int4(0, 0, 0, 0);
    bool   pred = false;
    if(i < np)
    {
        code = pairs[i];
        pred = friction_lag(x, code, dHat, kappa, &lam, c, B) >= 0;
    }
    unsigned k = wave_append(f.count, pred);
    if(pred && k < (unsigned)f.cap)
    {
        f.code[k]          = code;
        f.lam[k]           = lam;
        f.coord[2 * k]     = c[0];
        f.coord[2 * k + 1] = c[1];
        for(int q = 0; q < 6; q++)
            f.basis[6 * (size_t)k + q] = B[q];
    }
}

__global__ void __launch_bounds__(256) k_friction_ground_build(const unsigned* __restrict__ gcount,
                                                               const uint32_t* __restrict__ gp,
                                                               const double* __restrict__ x,
                                                               V3 n, double off, double dHat,
                                                               double kappa, FricDev f)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned ng = *gcount;
    if(i == 0)
        *f.gcount = ng;
    if(i >= ng)
        return;
    uint32_t v     = gp[i];
    double   dist  = dot(n, ld3(x + 3 * v)) - off;
    double   dist2 = dist * dist;
    double   t     = dist2 - dHat;
    double   g_b   = t * log(dist2 / dHat) * -2.0 - (t * t) / dist2;
    f.gids[i]      = v;
    f.glam[i]      = -kappa * 2.0 * sqrt(dist2) * g_b;
}

__global__ void __launch_bounds__(128) k_friction(const double* __restrict__ x, FricDev f,
                                                  BlockSink sink, double* __restrict__ gcontact,
                                                  int hessian)
{
    unsigned pi = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned nf = *f.count;
    if(pi >= (nf < (unsigned)f.cap ? nf : (unsigned)f.cap))
        return;
    int4          code = f.code[pi];
    const double* c    = f.coord + 2 * (size_t)pi;
    const double* B    = f.basis + 6 * (size_t)pi;
    int           ids[4];
    double        w[4], u[2];
    int           m   = rel_u(x, f.xprev, code, c, B, ids, w, u);
    double        eps = sqrt(f.eps2);
    double        s2  = u[0] * u[0] + u[1] * u[1], s = sqrt(s2);
    double        f1  = (s2 > f.eps2) ? 1.0 / s : f1_div_c1(s2, eps);
    double        sc  = f.mu * f.lam[pi];
    double        q0 = u[0] * f1, q1 = u[1] * f1;
    for(int k = 0; k < m; k++)
        for(int r = 0; r < 3; r++)
            atomicAdd(gcontact + 3 * ids[k] + r, sc * w[k] * (B[2 * r] * q0 + B[2 * r + 1] * q1));
    if(!hessian)
        return;
    double M2[4];
    if(s2 > f.eps2)
    {
        double c3 = 1.0 / (s2 * s);
        M2[0]     = 1.0 / s - u[0] * u[0] * c3;
        M2[1]     = -u[0] * u[1] * c3;
        M2[2]     = -u[1] * u[0] * c3;
        M2[3]     = 1.0 / s - u[1] * u[1] * c3;
    }
    else
    {
        double f2 = f2_c1(s2, eps);
        if(f2 != f1 && s2 != 0.0)
        {
            double c2 = (f1 - f2) / s2;
            M2[0]     = f1 - u[0] * u[0] * c2;
            M2[1]     = -u[0] * u[1] * c2;
            M2[2]     = -u[1] * u[0] * c2;
            M2[3]     = f1 - u[1] * u[1] * c2;
        }
        else
        {
            M2[0] = f1, M2[1] = 0.0, M2[2] = 0.0, M2[3] = f1;
        }
    }
    psd2(M2);
    // block (a, b) = sc w_a w_b Bb M2 Bb^T with Bb the 3x2 basis
    double P[9];
    for(int r = 0; r < 3; r++)
    {
        double a0 = B[2 * r] * M2[0] + B[2 * r + 1] * M2[2];
        double a1 = B[2 * r] * M2[1] + B[2 * r + 1] * M2[3];
        for(int cc = 0; cc < 3; cc++)
            P[3 * r + cc] = a0 * B[2 * cc] + a1 * B[2 * cc + 1];
    }
    // H_ab = w_a w_b (sc P)
    double K[9];
    for(int q = 0; q < 9; q++)
        K[q] = sc * P[q];
    sink.put_rank(ids, m, w, K);
}

__global__ void __launch_bounds__(256) k_friction_ground(const double* __restrict__ x, V3 n,
                                                         FricDev f, BlockSink sink,
                                                         const int* __restrict__ diag,
                                                         double* __restrict__ gcontact,
                                                         int hessian)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= *f.gcount)
        return;
    uint32_t v    = f.gids[i];
    V3       d    = ld3(x + 3 * v) - ld3(f.xprev + 3 * v);
    V3       vp   = d - n * dot(d, n);
    double   m2   = sqn(vp);
    double   eps  = sqrt(f.eps2);
    double   mult = f.mu_gd * f.glam[i];
    double   s    = (m2 > f.eps2) ? mult / sqrt(m2) : mult / eps;
    atomicAdd(gcontact + 3 * v + 0, vp.x * s);
    atomicAdd(gcontact + 3 * v + 1, vp.y * s);
    atomicAdd(gcontact + 3 * v + 2, vp.z * s);
    if(!hessian)
        return;
    double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if(m2 > f.eps2)
    {
        double mag  = sqrt(m2);
        double a    = vp.x * vp.x * -mult / m2 / mag + mult / mag;
        double b    = vp.x * vp.z * -mult / m2 / mag;
        double dd   = vp.z * vp.z * -mult / m2 / mag + mult / mag;
        double tr = 0.5 * (a + dd), df = 0.5 * (a - dd), r = sqrt(df * df + b * b);
        double Pm[4] = {0, 0, 0, 0};
        for(int e = 0; e < 2; e++)
        {
            double l = e ? tr + r : tr - r;
            if(!(l > 0.0))
                continue;
            double vx, vy;
            if(r == 0.0)
                vx = e ? 1.0 : 0.0, vy = e ? 0.0 : 1.0;
            else if((df >= 0.0) == (e == 1))
                vx = df + (e ? r : -r), vy = b;
            else
                vx = b, vy = (e ? r : -r) - df;
            double nv = sqrt(vx * vx + vy * vy);
            if(nv == 0.0)
                continue;
            vx /= nv, vy /= nv;
            Pm[0] += l * vx * vx;
            Pm[1] += l * vx * vy;
            Pm[2] += l * vx * vy;
            Pm[3] += l * vy * vy;
        }
        H[0] = Pm[0], H[2] = Pm[1], H[6] = Pm[2], H[8] = Pm[3];
    }
    else
    {
        H[0] = mult / eps;
        H[8] = mult / eps;
    }
    (void)diag;
    sink.put((int)v, (int)v, H, true);
}

__global__ void __launch_bounds__(256) k_energy_friction(const double* __restrict__ x, FricDev f,
                                                         double* out)
{
    unsigned i  = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned nf = *f.count;
    double   e  = 0.0;
    if(i < (nf < (unsigned)f.cap ? nf : (unsigned)f.cap))
    {
        int    ids[4];
        double w[4], u[2];
        rel_u(x, f.xprev, f.code[i], f.coord + 2 * (size_t)i, f.basis + 6 * (size_t)i, ids, w,
              u);
        double s2 = u[0] * u[0] + u[1] * u[1];
        e = f.mu * (s2 > f.eps2 ? f.lam[i] * sqrt(s2) : f.lam[i] * f0_c1(s2, sqrt(f.eps2)));
    }
    block_partial(e, out);
}

__global__ void __launch_bounds__(256) k_energy_friction_ground(const double* __restrict__ x, V3 n,
                                                                FricDev f, double* out)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    double   e = 0.0;
    if(i < *f.gcount)
    {
        uint32_t v   = f.gids[i];
        V3       d   = ld3(x + 3 * v) - ld3(f.xprev + 3 * v);
        V3       vp  = d - n * dot(d, n);
        double   m2  = sqn(vp);
        double   eps = sqrt(f.eps2);
        e = f.mu_gd * (m2 > eps * eps ? f.glam[i] * (sqrt(m2) - eps * 0.5)
                                      : f.glam[i] * m2 / eps * 0.5);
    }
    block_partial(e, out);
}

// ------------------------------------------------------------- launchers
void launch_friction_build(const SimDev& d, double kappa, hipStream_t s)
{
    const FricDev& f = d.fric;
    SG_HIP(hipMemsetAsync(f.count, 0, sizeof(unsigned), s));
    SG_HIP(hipMemsetAsync(f.gcount, 0, sizeof(unsigned), s));
    if(f.mu != 0.0 && d.pair_cap > 0)
        hipLaunchKernelGGL(k_friction_build, dim3(grid_for(d.pair_cap, 256)), dim3(256), 0, s,
                           d.pcount, d.pair_cap, d.pairs, d.x, d.dHat, kappa, f);
    if(f.mu_gd != 0.0 && d.use_ground && d.ns > 0)
        hipLaunchKernelGGL(k_friction_ground_build, dim3(grid_for(d.ns, 256)), dim3(256), 0, s,
                           d.gcount, d.gp, d.x, d.gn, d.goff, d.dHat, kappa, f);
    SG_LAUNCH_CHECK();
}

void launch_friction(const SimDev& d, bool hessian, double* gcontact, hipStream_t s)
{
    const FricDev& f = d.fric;
    if(f.mu != 0.0 && f.cap > 0)
        hipLaunchKernelGGL(k_friction, dim3(grid_for(f.cap, 128)), dim3(128), 0, s, d.x, f,
                           block_sink(d), gcontact, hessian ? 1 : 0);
    if(f.mu_gd != 0.0 && d.use_ground && d.ns > 0)
        hipLaunchKernelGGL(k_friction_ground, dim3(grid_for(d.ns, 256)), dim3(256), 0, s, d.x,
                           d.gn, f, block_sink(d), d.diag, gcontact, hessian ? 1 : 0);
    SG_LAUNCH_CHECK();
}

void launch_energy_friction(const SimDev& d, const EnergyLayout& L, double* partial, hipStream_t s)
{
    const FricDev& f = d.fric;
    if(L.n[4])
        hipLaunchKernelGGL(k_energy_friction, dim3(L.n[4]), dim3(256), 0, s, d.x, f,
                           partial + L.off[4]);
    if(L.n[5])
        hipLaunchKernelGGL(k_energy_friction_ground, dim3(L.n[5]), dim3(256), 0, s, d.x, d.gn, f,
                           partial + L.off[5]);
    SG_LAUNCH_CHECK();
}

}  // namespace sgipc
