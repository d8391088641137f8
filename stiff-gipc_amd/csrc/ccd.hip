// ccd.hip -- step-size bounds: additive CCD over candidate pairs, the ground
// bound and the CFL speed.
//
//   k_self_ccd    1 / toc per candidate, max-reduced
//                 (_reduct_min_selfTimeStep_to_double GIPC.cu:7054-7157 with
//                 point_triangle_ccd / edge_edge_ccd, ACCD.cu:373-524)
//   k_ground_ccd  (n.dx) / (dist * slackness), max-reduced (GIPC.cu:6904-6980)
//   k_cfl         max |dx| over surface vertices (GIPC.cu:7159-7217)
// The host takes 1 / max.  Maxima of positive doubles are reduced with a
// 64-bit integer atomicMax on the bit pattern after a wave64 max.
// Built with -ffp-contract=off (the ACCD iteration is order sensitive).
#include "sim_kernels.h"

namespace sgipc
{

__device__ double pt_ccd(V3 p, V3 t0, V3 t1, V3 t2, V3 dp, V3 dt0, V3 dt1, V3 dt2, double eta,
                         double th)
{
    V3 mov = ((dt0 + dt1) + (dt2 + dp)) * -0.25;
    dt0    = dt0 + mov;
    dt1    = dt1 + mov;
    dt2    = dt2 + mov;
    dp     = dp + mov;
    double m0 = sqn(dt0), m1 = sqn(dt1), m2 = sqn(dt2);
    double mm = (m1 < m2) ? m2 : m1;
    mm        = (m0 < mm) ? mm : m0;
    double maxd = sqrt(sqn(dp)) + sqrt(mm);
    if(maxd == 0)
        return 1.0;
    double d2  = pt_dist_unclassified(p, t0, t1, t2);
    double d   = sqrt(d2);
    double gap = eta * (d2 - th * th) / (d + th);
    double toc = 0.0;
    int    cnt = 0;
    while(true)
    {
        cnt++;
        if(cnt > 50000)
            return toc;
        double lb = (1 - eta) * (d2 - th * th) / ((d + th) * maxd);
        p         = p + dp * lb;
        t0        = t0 + dt0 * lb;
        t1        = t1 + dt1 * lb;
        t2        = t2 + dt2 * lb;
        d2        = pt_dist_unclassified(p, t0, t1, t2);
        d         = sqrt(d2);
        if(toc && ((d2 - th * th) / (d + th) < gap))
            break;
        toc += lb;
        if(toc > 1.0)
            return 1.0;
    }
    return toc;
}

__device__ __forceinline__ double min4_pp(V3 a0, V3 a1, V3 b0, V3 b1)
{
    double d0 = sqn(a0 - b0), d1 = sqn(a0 - b1), d2 = sqn(a1 - b0), d3 = sqn(a1 - b1);
    double x  = (d1 < d0) ? d1 : d0;
    double y  = (d3 < d2) ? d3 : d2;
    return (y < x) ? y : x;
}

__device__ double ee_ccd(V3 a0, V3 a1, V3 b0, V3 b1, V3 da0, V3 da1, V3 db0, V3 db1, double eta,
                         double th)
{
    V3 mov = ((da0 + da1) + (db0 + db1)) * -0.25;
    da0    = da0 + mov;
    da1    = da1 + mov;
    db0    = db0 + mov;
    db1    = db1 + mov;
    double qa = sqn(da0), qb = sqn(da1), qc = sqn(db0), qd = sqn(db1);
    double maxd = sqrt((qa < qb) ? qb : qa) + sqrt((qc < qd) ? qd : qc);
    if(maxd == 0)
        return 1.0;
    double d2 = ee_dist_unclassified(a0, a1, b0, b1);
    double dF = d2 - th * th;
    if(dF <= 0)
    {
        d2 = min4_pp(a0, a1, b0, b1);
        dF = d2 - th * th;
    }
    double d   = sqrt(d2);
    double gap = eta * dF / (d + th);
    double toc = 0.0;
    int    cnt = 0;
    while(true)
    {
        cnt++;
        if(cnt > 50000)
            return toc;
        double lb = (1 - eta) * dF / ((d + th) * maxd);
        a0        = a0 + da0 * lb;
        a1        = a1 + da1 * lb;
        b0        = b0 + db0 * lb;
        b1        = b1 + db1 * lb;
        d2        = ee_dist_unclassified(a0, a1, b0, b1);
        dF        = d2 - th * th;
        if(dF <= 0)
        {
            d2 = min4_pp(a0, a1, b0, b1);
            dF = d2 - th * th;
        }
        d = sqrt(d2);
        if(toc && (dF / (d + th) < gap))
            break;
        toc += lb;
        if(toc > 1.0)
            return 1.0;
    }
    return toc;
}

__global__ void __launch_bounds__(256) k_self_ccd(const double* __restrict__ x,
                                                  const double* __restrict__ dx,
                                                  const int4* __restrict__ cand,
                                                  const unsigned* __restrict__ count, int cap,
                                                  double slackness, double* out)
{
    unsigned i    = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned n    = *count;
    n             = n < (unsigned)cap ? n : (unsigned)cap;
    double   temp = 0.0;
    if(i < n)
    {
        int4   c   = cand[i];
        double eta = 1.0 - slackness;
        if(c.x < 0)
        {
            int p = -c.x - 1;
            temp  = 1.0
                   / pt_ccd(ld3(x + 3 * p), ld3(x + 3 * c.y), ld3(x + 3 * c.z), ld3(x + 3 * c.w),
                            ld3(dx + 3 * p) * -1, ld3(dx + 3 * c.y) * -1, ld3(dx + 3 * c.z) * -1,
                            ld3(dx + 3 * c.w) * -1, eta, 0);
        }
        else
            temp = 1.0
                   / ee_ccd(ld3(x + 3 * c.x), ld3(x + 3 * c.y), ld3(x + 3 * c.z),
                            ld3(x + 3 * c.w), ld3(dx + 3 * c.x) * -1, ld3(dx + 3 * c.y) * -1,
                            ld3(dx + 3 * c.z) * -1, ld3(dx + 3 * c.w) * -1, eta, 0);
    }
    temp = wave_max(temp);
    if(__lane_id() == 0 && temp > 0.0)
        atomic_max_pos(out, temp);
}

__global__ void __launch_bounds__(256) k_ground_ccd(int ns, const uint32_t* __restrict__ surf,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ dx, V3 n,
                                                    double off, double slackness, double* out)
{
    int    i    = blockIdx.x * blockDim.x + threadIdx.x;
    double temp = 0.0;
    if(i < ns)
    {
        int    v    = (int)surf[i];
        double coef = dot(n, ld3(dx + 3 * v));
        temp        = 1.0;
        if(coef > 0.0)
        {
            double dist = dot(n, ld3(x + 3 * v)) - off;
            temp        = coef / (dist * slackness);
        }
    }
    temp = wave_max(temp);
    if(__lane_id() == 0 && temp > 0.0)
        atomic_max_pos(out, temp);
}

__global__ void __launch_bounds__(256) k_cfl(int ns, const uint32_t* __restrict__ surf,
                                             const double* __restrict__ dx, double* out)
{
    int    i = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if(i < ns)
        s = sqrt(sqn(ld3(dx + 3 * surf[i])));
    s = wave_max(s);
    if(__lane_id() == 0 && s > 0.0)
        atomic_max_pos(out, s);
}

void launch_self_ccd(const double* x, const double* dx, const int4* cand, const unsigned* count,
                     int cap, double slackness, double* out_max, hipStream_t s)
{
    if(cap <= 0)
        return;
    hipLaunchKernelGGL(k_self_ccd, dim3(grid_for(cap, 256)), dim3(256), 0, s, x, dx, cand, count,
                       cap, slackness, out_max);
    SG_LAUNCH_CHECK();
}

void launch_ground_ccd(const SimDev& d, const double* dx, double slackness, double* out_max,
                       hipStream_t s)
{
    if(d.ns == 0)
        return;
    hipLaunchKernelGGL(k_ground_ccd, dim3(grid_for(d.ns, 256)), dim3(256), 0, s, d.ns, d.surf, d.x,
                       dx, d.gn, d.goff, slackness, out_max);
    SG_LAUNCH_CHECK();
}

void launch_cfl(const SimDev& d, const double* dx, double* out_max, hipStream_t s)
{
    if(d.ns == 0)
        return;
    hipLaunchKernelGGL(k_cfl, dim3(grid_for(d.ns, 256)), dim3(256), 0, s, d.ns, d.surf, dx,
                       out_max);
    SG_LAUNCH_CHECK();
}

}  // namespace sgipc
