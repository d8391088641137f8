// contact.hip -- proximity and CCD candidate queries on the LBVHs.
//
//   k_query_vf   surface vertex vs face tree  (_selfQuery_vf, mlbvh.cu:1308-1408)
//   k_query_ee   edge leaf vs edge tree       (_selfQuery_ee, mlbvh.cu:1511-1634)
//   k_ccd_vf/ee  swept candidates             (_selfQuery_*_ccd, mlbvh.cu:1410-1740)
//   k_edge_tri   segment-triangle intersection (_edgeTriIntersectionQuery,
//                GIPC.cu:8188-8302, segTriIntersect :961-1013)
// Classification and codes follow _checkPTintersection / _checkEEintersection
// (mlbvh.cu:537-1062, smooth == false).  Built with -ffp-contract=off so the
// pair multiset equals the CPU restatement's.
//
// One wave per 64 queries; the traversal stack lives in LDS as stk[depth][lane]
// (64 levels, the depth bound of a 64-bit-key Karras tree), so pushes and pops
// are bank-conflict free.  Each record carries its raw candidate
// (_ccd_collisionPair) for the CCD step bound.
#include "sim_kernels.h"

namespace sgipc
{

constexpr int kStack = 64;

__device__ __forceinline__ bool overlap(const double* l, const double* r, double gap)
{
    // l, r: {upper xyz, lower xyz} (mlbvh.cu:31-40)
    for(int c = 0; c < 3; c++)
        if((r[3 + c] - l[c]) >= gap || (l[3 + c] - r[c]) >= gap)
            return false;
    return true;
}

struct PairSink
{
    int4*     pairs;
    int4*     raw;
    unsigned* cnt;  // [0] total, [2] PP, [3] PE, [4] PT + EE
    int       cap;
    int*      overflow;
    __device__ void emit(int4 code, int4 rawc, int slot_type)
    {
        unsigned o = atomicAdd(cnt, 1u);
        if((int)o >= cap)
        {
            atomicExch(overflow, 1);
            return;
        }
        pairs[o] = code;
        raw[o]   = rawc;
        atomicAdd(cnt + slot_type, 1u);
    }
};

__device__ void check_pt(const double* x, int id0, int id1, int id2, int id3, double dHat,
                         PairSink& sk)
{
    V3   v0 = ld3(x + 3 * id0), v1 = ld3(x + 3 * id1), v2 = ld3(x + 3 * id2), v3 = ld3(x + 3 * id3);
    int  t  = dtype_pt(v0, v1, v2, v3);
    int4 raw = make_int4(-id0 - 1, id1, id2, id3);
    switch(t)
    {
        case 0:
            if(d_pp(v0, v1) < dHat)
                sk.emit(make_int4(-id0 - 1, id1, -1, -1), raw, 2);
            break;
        case 1:
            if(d_pp(v0, v2) < dHat)
                sk.emit(make_int4(-id0 - 1, id2, -1, -1), raw, 2);
            break;
        case 2:
            if(d_pp(v0, v3) < dHat)
                sk.emit(make_int4(-id0 - 1, id3, -1, -1), raw, 2);
            break;
        case 3:
            if(d_pe(v0, v1, v2) < dHat)
                sk.emit(make_int4(-id0 - 1, id1, id2, -1), raw, 3);
            break;
        case 4:
            if(d_pe(v0, v2, v3) < dHat)
                sk.emit(make_int4(-id0 - 1, id2, id3, -1), raw, 3);
            break;
        case 5:
            if(d_pe(v0, v3, v1) < dHat)
                sk.emit(make_int4(-id0 - 1, id3, id1, -1), raw, 3);
            break;
        case 6:
            if(d_pt(v0, v1, v2, v3) < dHat)
                sk.emit(make_int4(-id0 - 1, id1, id2, id3), raw, 4);
            break;
        default: break;
    }
}

__device__ __forceinline__ double epx(V3 r0, V3 r1, V3 r2, V3 r3)
{
    return 1e-3 * sqn(r0 - r1) * sqn(r2 - r3);
}

__device__ void check_ee(const double* x, const double* xr, int id0, int id1, int id2, int id3,
                         int obj, double dHat, PairSink& sk)
{
    V3   v0 = ld3(x + 3 * id0), v1 = ld3(x + 3 * id1), v2 = ld3(x + 3 * id2), v3 = ld3(x + 3 * id3);
    int  t   = dtype_ee(v0, v1, v2, v3);
    int4 raw = make_int4(id0, id1, id2, id3);
    auto flag01 = [&]() {
        double cr = sqn(cross(v0 - v1, v2 - v3));
        return (cr < epx(ld3(xr + 3 * id0), ld3(xr + 3 * id1), ld3(xr + 3 * id2),
                         ld3(xr + 3 * id3)))
                   ? -obj - 2
                   : -1;
    };
    auto flag23 = [&]() {
        double cr = sqn(cross(v2 - v3, v0 - v1));
        return (cr < epx(ld3(xr + 3 * id2), ld3(xr + 3 * id3), ld3(xr + 3 * id0),
                         ld3(xr + 3 * id1)))
                   ? -obj - 2
                   : -1;
    };
    switch(t)
    {
        case 0:
            if(d_pp(v0, v2) < dHat)
                sk.emit(make_int4(-id0 - 1, id2, -1, flag01()), raw, 2);
            break;
        case 1:
            if(d_pp(v0, v3) < dHat)
                sk.emit(make_int4(-id0 - 1, id3, -1, flag01()), raw, 2);
            break;
        case 2:
            if(d_pe(v0, v2, v3) < dHat)
                sk.emit(make_int4(-id0 - 1, id2, id3, flag01()), raw, 3);
            break;
        case 3:
            if(d_pp(v1, v2) < dHat)
                sk.emit(make_int4(-id1 - 1, id2, -1, flag01()), raw, 2);
            break;
        case 4:
            if(d_pp(v1, v3) < dHat)
                sk.emit(make_int4(-id1 - 1, id3, -1, flag01()), raw, 2);
            break;
        case 5:
            if(d_pe(v1, v2, v3) < dHat)
                sk.emit(make_int4(-id1 - 1, id2, id3, flag01()), raw, 3);
            break;
        case 6:
            if(d_pe(v2, v0, v1) < dHat)
                sk.emit(make_int4(-id2 - 1, id0, id1, flag23()), raw, 3);
            break;
        case 7:
            if(d_pe(v3, v0, v1) < dHat)
                sk.emit(make_int4(-id3 - 1, id0, id1, flag23()), raw, 3);
            break;
        case 8:
            if(d_ee(v0, v1, v2, v3) < dHat)
                sk.emit(make_int4(id0, id1, id2, id3), raw, 4);
            break;
        default: break;
    }
}

__device__ __forceinline__ bool vf_ok(const int* body, const int* bt, int v, uint3 f)
{
    if(!((body[v] != body[f.x]) || (body[v] == -1)))
        return false;
    if(v == (int)f.x || v == (int)f.y || v == (int)f.z)
        return false;
    if(bt[v] >= 2 && bt[f.x] >= 2 && bt[f.y] >= 2 && bt[f.z] >= 2)
        return false;
    return true;
}

__device__ __forceinline__ bool ee_ok(const int* body, const int* bt, uint2 a, uint2 b, int self,
                                      int obj)
{
    if(self == obj)
        return false;
    if(!((body[a.x] != body[b.x]) || (body[a.x] == -1)))
        return false;
    if(a.x == b.x || a.x == b.y || a.y == b.x || a.y == b.y || obj < self)
        return false;
    if(bt[a.x] >= 2 && bt[a.y] >= 2 && bt[b.x] >= 2 && bt[b.y] >= 2)
        return false;
    return true;
}

// generic stack traversal: visit(leaf_element) for every overlapping leaf.
// With rng (each internal node's sorted leaf range) only leaves at sorted
// positions above `after` are visited: a self query of the tree from its
// sorted leaf `after` then meets every unordered pair of leaves once, from the
// earlier one, and skips the subtrees that lie wholly at or before it.
template <typename Visit>
__device__ __forceinline__ void traverse(const uint32_t* __restrict__ nodes,
                                         const double* __restrict__ bvs, int n, const double* q,
                                         double gap, uint32_t (*stk)[64], Visit&& visit,
                                         const int* __restrict__ rng = nullptr, int after = -1)
{
    const int lane = threadIdx.x;
    if(n == 1)
    {
        if(overlap(q, bvs, gap))
            visit(nodes[3]);
        return;
    }
    int sp          = 0;
    stk[sp++][lane] = 0;
    do
    {
        uint32_t node = stk[--sp][lane];
        uint32_t ch[2] = {nodes[4 * (size_t)node + 1], nodes[4 * (size_t)node + 2]};
        for(int c = 0; c < 2; c++)
        {
            uint32_t C = ch[c];
            if(rng)
            {
                const int last = (int)C >= n - 1 ? (int)C - (n - 1) : rng[2 * (size_t)C + 1];
                if(last <= after)
                    continue;
            }
            if(!overlap(q, bvs + 6 * (size_t)C, gap))
                continue;
            uint32_t obj = nodes[4 * (size_t)C + 3];
            if(obj != 0xFFFFFFFFu)
                visit(obj);
            else if(sp < kStack)
                stk[sp++][lane] = C;
        }
    } while(sp > 0);
}

__global__ void __launch_bounds__(64) k_query_vf(int ns, const uint32_t* __restrict__ surf,
                                                 const uint3* __restrict__ faces,
                                                 const int* __restrict__ body,
                                                 const int* __restrict__ bt,
                                                 const double* __restrict__ x,
                                                 const uint32_t* __restrict__ nodes,
                                                 const double* __restrict__ bvs, int nf,
                                                 double dHat, PairSink sk)
{
    __shared__ uint32_t stk[kStack][64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= ns)
        return;
    int    v    = (int)surf[i];
    double q[6] = {x[3 * v], x[3 * v + 1], x[3 * v + 2], x[3 * v], x[3 * v + 1], x[3 * v + 2]};
    double gap  = sqrt(dHat);
    traverse(nodes, bvs, nf, q, gap, stk, [&](uint32_t obj) {
        uint3 f = faces[obj];
        if(vf_ok(body, bt, v, f))
            check_pt(x, v, f.x, f.y, f.z, dHat, sk);
    });
}

// Edge-edge pairs: each edge queries the edge tree from its own sorted leaf
// for the leaves after it, and a pair is emitted with its lower edge index
// first, as the reference emits it from that edge's query (obj > self).
__global__ void __launch_bounds__(64) k_query_ee(int ne, const uint2* __restrict__ edges,
                                                 const int* __restrict__ body,
                                                 const int* __restrict__ bt,
                                                 const double* __restrict__ x,
                                                 const double* __restrict__ xr,
                                                 const uint32_t* __restrict__ nodes,
                                                 const double* __restrict__ bvs,
                                                 const int* __restrict__ rng, double dHat,
                                                 PairSink sk)
{
    __shared__ uint32_t stk[kStack][64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= ne)
        return;
    int    leaf = i + ne - 1;
    int    self = (int)nodes[4 * (size_t)leaf + 3];
    double q[6];
    for(int c = 0; c < 6; c++)
        q[c] = bvs[6 * (size_t)leaf + c];
    uint2  e   = edges[self];
    double gap = sqrt(dHat);
    traverse(
        nodes, bvs, ne, q, gap, stk,
        [&](uint32_t obj) {
            const int lo = self < (int)obj ? self : (int)obj;
            const int hi = self < (int)obj ? (int)obj : self;
            uint2     o  = edges[obj];
            uint2     a  = lo == self ? e : o, b = lo == self ? o : e;
            if(ee_ok(body, bt, a, b, lo, hi))
                check_ee(x, xr, a.x, a.y, b.x, b.y, hi, dHat, sk);
        },
        rng, i);
}

struct CandSink
{
    int4*     out;
    unsigned* cnt;
    int       cap;
    int*      overflow;
    __device__ void emit(int4 c)
    {
        unsigned o = atomicAdd(cnt, 1u);
        if((int)o >= cap)
        {
            atomicExch(overflow, 1);
            return;
        }
        out[o] = c;
    }
};

__global__ void __launch_bounds__(64) k_ccd_vf(int ns, const uint32_t* __restrict__ surf,
                                               const uint3* __restrict__ faces,
                                               const int* __restrict__ body,
                                               const int* __restrict__ bt,
                                               const double* __restrict__ x,
                                               const double* __restrict__ dx, double alpha,
                                               const uint32_t* __restrict__ nodes,
                                               const double* __restrict__ bvs, int nf,
                                               double dHat, CandSink sk)
{
    __shared__ uint32_t stk[kStack][64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= ns)
        return;
    int    v = (int)surf[i];
    double q[6];
    for(int c = 0; c < 3; c++)
    {
        double a = x[3 * v + c], b = x[3 * v + c] - dx[3 * v + c] * alpha;
        q[c]     = (a < b) ? b : a;   // upper: combines(v - dx alpha) after (v, v)
        q[3 + c] = (b < a) ? b : a;
    }
    double gap = sqrt(dHat);
    traverse(nodes, bvs, nf, q, gap, stk, [&](uint32_t obj) {
        uint3 f = faces[obj];
        if(((body[v] != body[f.x]) || (body[v] == -1))
           && !(bt[v] >= 2 && bt[f.x] >= 2 && bt[f.y] >= 2 && bt[f.z] >= 2)
           && v != (int)f.x && v != (int)f.y && v != (int)f.z)
            sk.emit(make_int4(-v - 1, f.x, f.y, f.z));
    });
}

__global__ void __launch_bounds__(64) k_ccd_ee(int ne, const uint2* __restrict__ edges,
                                               const int* __restrict__ body,
                                               const int* __restrict__ bt,
                                               const uint32_t* __restrict__ nodes,
                                               const double* __restrict__ bvs,
                                               const int* __restrict__ rng, double dHat,
                                               CandSink sk)
{
    __shared__ uint32_t stk[kStack][64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= ne)
        return;
    int    leaf = i + ne - 1;
    int    self = (int)nodes[4 * (size_t)leaf + 3];
    double q[6];
    for(int c = 0; c < 6; c++)
        q[c] = bvs[6 * (size_t)leaf + c];
    uint2  e   = edges[self];
    double gap = sqrt(dHat);
    traverse(
        nodes, bvs, ne, q, gap, stk,
        [&](uint32_t obj) {
            const int lo = self < (int)obj ? self : (int)obj;
            const int hi = self < (int)obj ? (int)obj : self;
            uint2     o  = edges[obj];
            uint2     a  = lo == self ? e : o, b = lo == self ? o : e;
            if(ee_ok(body, bt, a, b, lo, hi))
                sk.emit(make_int4(a.x, a.y, b.x, b.y));
        },
        rng, i);
}

__device__ bool seg_tri(V3 e0, V3 e1, V3 t0, V3 t1, V3 t2)
{
    V3 c0 = t1 - t0, c1 = t2 - t0, c2 = e0 - e1;
    V3 n  = cross(c0, c1);
    if(dot(n, e0 - t0) * dot(n, e1 - t0) > 0)
        return false;
    double det = det_cols(c0, c1, c2);
    // segTriIntersect's |det| < 1e-20, plus a scale-relative test so that an
    // edge lying in the triangle's plane (coplanar surface faces) is rejected
    // whatever the mesh size; the absolute constant alone lets round-off
    // determinants through at centimetre scales
    double sc = sqrt(sqn(c0) * sqn(c1) * sqn(c2));
    if(fabs(det) < 1e-20 || fabs(det) <= 1e-12 * sc)
        return false;
    V3     b = e0 - t0;
    double u = det_cols(b, c1, c2) / det, v = det_cols(c0, b, c2) / det,
           t = det_cols(c0, c1, b) / det;
    return (u >= 0.0 && v >= 0.0 && u + v <= 1.0 && t >= 0.0 && t <= 1.0);
}

__global__ void __launch_bounds__(64) k_edge_tri(int nf, const uint3* __restrict__ faces,
                                                 const uint2* __restrict__ edges,
                                                 const int* __restrict__ bt,
                                                 const double* __restrict__ x,
                                                 const uint32_t* __restrict__ nodes,
                                                 const double* __restrict__ bvs, int ne, int* hit)
{
    __shared__ uint32_t stk[kStack][64];
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if(f >= nf)
        return;
    uint3  F = faces[f];
    double q[6] = {-1e32, -1e32, -1e32, 1e32, 1e32, 1e32};
    uint32_t vs[3] = {F.x, F.y, F.z};
    for(int k = 0; k < 3; k++)
        for(int c = 0; c < 3; c++)
        {
            double p = x[3 * vs[k] + c];
            q[3 + c] = (p < q[3 + c]) ? p : q[3 + c];
            q[c]     = (q[c] < p) ? p : q[c];
        }
    bool found = false;
    traverse(nodes, bvs, ne, q, 0.0, stk, [&](uint32_t obj) {
        if(found)
            return;
        uint2 E = edges[obj];
        if(F.x == E.x || F.x == E.y || F.y == E.x || F.y == E.y || F.z == E.x || F.z == E.y)
            return;
        if(bt[F.x] >= 2 && bt[F.y] >= 2 && bt[F.z] >= 2 && bt[E.x] >= 2 && bt[E.y] >= 2)
            return;
        if(seg_tri(ld3(x + 3 * E.x), ld3(x + 3 * E.y), ld3(x + 3 * F.x), ld3(x + 3 * F.y),
                   ld3(x + 3 * F.z)))
            found = true;
    });
    if(found)
        atomicExch(hit, 1);
}

__global__ void k_ground_intersect(const unsigned* __restrict__ gcount,
                                   const uint32_t* __restrict__ gp, const double* __restrict__ x,
                                   V3 n, double off, int* hit)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= *gcount)
        return;
    double dist = dot(n, ld3(x + 3 * gp[i])) - off;
    if(dist < 0)
        atomicExch(hit, 1);
}

// ----------------------------------------------------------------- launch
void launch_query_proximity(const SimDev& d, const Lbvh& bf, const Lbvh& be, const double* xr,
                            hipStream_t s)
{
    SG_HIP(hipMemsetAsync(d.pcount, 0, 5 * sizeof(unsigned), s));
    PairSink sk{d.pairs, d.pairs_raw, d.pcount, d.pair_cap, d.overflow};
    if(d.nf > 0 && d.ns > 0)
        hipLaunchKernelGGL(k_query_vf, dim3(grid_for(d.ns, 64)), dim3(64), 0, s, d.ns, d.surf,
                           (const uint3*)d.faces, d.body, d.btype, d.x, bf.nodes.data(),
                           bf.bvs.data(), d.nf, d.dHat, sk);
    if(d.ne > 1)
        hipLaunchKernelGGL(k_query_ee, dim3(grid_for(d.ne, 64)), dim3(64), 0, s, d.ne,
                           (const uint2*)d.edges, d.body, d.btype, d.x, xr, be.nodes.data(),
                           be.bvs.data(), be.rng.data(), d.dHat, sk);
    SG_LAUNCH_CHECK();
}

void launch_query_ccd(const SimDev& d, const Lbvh& bf, const Lbvh& be, const double* dx,
                      double alpha, int4* cand, unsigned* ccount, int cap, int* overflow,
                      hipStream_t s)
{
    SG_HIP(hipMemsetAsync(ccount, 0, sizeof(unsigned), s));
    CandSink sk{cand, ccount, cap, overflow};
    if(d.nf > 0 && d.ns > 0)
        hipLaunchKernelGGL(k_ccd_vf, dim3(grid_for(d.ns, 64)), dim3(64), 0, s, d.ns, d.surf,
                           (const uint3*)d.faces, d.body, d.btype, d.x, dx, alpha,
                           bf.nodes.data(), bf.bvs.data(), d.nf, d.dHat, sk);
    if(d.ne > 1)
        hipLaunchKernelGGL(k_ccd_ee, dim3(grid_for(d.ne, 64)), dim3(64), 0, s, d.ne,
                           (const uint2*)d.edges, d.body, d.btype, be.nodes.data(), be.bvs.data(),
                           be.rng.data(), d.dHat, sk);
    SG_LAUNCH_CHECK();
}

void launch_edge_tri(const SimDev& d, const Lbvh& be, int* hit, hipStream_t s)
{
    if(d.nf == 0 || d.ne == 0)
        return;
    hipLaunchKernelGGL(k_edge_tri, dim3(grid_for(d.nf, 64)), dim3(64), 0, s, d.nf,
                       (const uint3*)d.faces, (const uint2*)d.edges, d.btype, d.x,
                       be.nodes.data(), be.bvs.data(), d.ne, hit);
    SG_LAUNCH_CHECK();
}

void launch_ground_intersect(const SimDev& d, int* hit, hipStream_t s)
{
    if(!d.use_ground || d.ns == 0)
        return;
    hipLaunchKernelGGL(k_ground_intersect, dim3(grid_for(d.ns, 256)), dim3(256), 0, s, d.gcount,
                       d.gp, d.x, d.gn, d.goff, hit);
    SG_LAUNCH_CHECK();
}

}  // namespace sgipc
