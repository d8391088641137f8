// lbvh.hip -- linear BVH over collision faces / edges (lbvh_f / lbvh_e).
//
// Construction follows StiffGIPC/mlbvh.cu:2074-2157 (Karras 2012):
//   leaf AABBs (:1149-1204), scene box (calcMaxBV :1745), keys
//   morton30(centre) << 32 | index (:1208-1227), sort, leaf/internal nodes
//   (:1229-1269, determine_range :135, find_split :198), bottom-up refit (:1271).
// Built with -ffp-contract=off: keys, node arrays and boxes are bit-identical
// to the CPU restatement.
//
// The key sort is this file's own LSD radix sort on 8-bit digits (keys only:
// the primitive index is the key's low word).  Each 256-thread block ranks
// its 2048-key tile stably with wave64 multi-split: eight ballots give every
// lane the mask of same-digit peers, per-wave digit counts are combined in
// LDS.  Digits above the index width and above the 30-bit Morton code are
// skipped, so a 50k-primitive tree needs 6 passes.
#include "sim_kernels.h"

namespace sgipc
{

constexpr int kSortThreads = 256;
constexpr int kSortItems   = 8;
constexpr int kSortTile    = kSortThreads * kSortItems;

__global__ void k_leaf_aabb(int n, int arity, const uint32_t* __restrict__ prims,
                            const double* __restrict__ x, const double* __restrict__ dx,
                            double alpha, double* __restrict__ leaf)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if(e >= n)
        return;
    double up[3] = {-1e32, -1e32, -1e32}, lo[3] = {1e32, 1e32, 1e32};
    for(int k = 0; k < arity; k++)
    {
        uint32_t v = prims[(size_t)arity * e + k];
        for(int c = 0; c < 3; c++)
        {
            double p = x[3 * (size_t)v + c];
            lo[c]    = (p < lo[c]) ? p : lo[c];
            up[c]    = (up[c] < p) ? p : up[c];
            if(dx)
            {
                double q = p - dx[3 * (size_t)v + c] * alpha;
                lo[c]    = (q < lo[c]) ? q : lo[c];
                up[c]    = (up[c] < q) ? q : up[c];
            }
        }
    }
    for(int c = 0; c < 3; c++)
    {
        leaf[6 * (size_t)e + c]     = up[c];
        leaf[6 * (size_t)e + 3 + c] = lo[c];
    }
}

// min/max reduction of the leaf boxes: per-block partials then one block
__global__ void __launch_bounds__(256) k_box_partial(int n, const double* __restrict__ leaf,
                                                     double* __restrict__ part)
{
    __shared__ double s[6][4];
    int    e = blockIdx.x * blockDim.x + threadIdx.x;
    double v[6];
    for(int c = 0; c < 3; c++)
    {
        v[c]     = (e < n) ? leaf[6 * (size_t)e + c] : -1e32;
        v[3 + c] = (e < n) ? leaf[6 * (size_t)e + 3 + c] : 1e32;
    }
    for(int c = 0; c < 3; c++)
    {
        v[c]     = wave_max(v[c]);
        v[3 + c] = wave_min(v[3 + c]);
    }
    if(__lane_id() == 0)
        for(int c = 0; c < 6; c++)
            s[c][threadIdx.x >> 6] = v[c];
    __syncthreads();
    if(threadIdx.x < 6)
    {
        int    c = threadIdx.x;
        double r = s[c][0];
        for(int w = 1; w < 4; w++)
            r = (c < 3) ? ((r < s[c][w]) ? s[c][w] : r) : ((s[c][w] < r) ? s[c][w] : r);
        part[6 * (size_t)blockIdx.x + c] = r;
    }
}

__global__ void __launch_bounds__(256) k_box_final(int nb, const double* __restrict__ part,
                                                   double* __restrict__ scene)
{
    __shared__ double s[6][4];
    double v[6] = {-1e32, -1e32, -1e32, 1e32, 1e32, 1e32};
    for(int i = threadIdx.x; i < nb; i += blockDim.x)
        for(int c = 0; c < 6; c++)
        {
            double p = part[6 * (size_t)i + c];
            v[c]     = (c < 3) ? ((v[c] < p) ? p : v[c]) : ((p < v[c]) ? p : v[c]);
        }
    for(int c = 0; c < 3; c++)
    {
        v[c]     = wave_max(v[c]);
        v[3 + c] = wave_min(v[3 + c]);
    }
    if(__lane_id() == 0)
        for(int c = 0; c < 6; c++)
            s[c][threadIdx.x >> 6] = v[c];
    __syncthreads();
    if(threadIdx.x < 6)
    {
        int    c = threadIdx.x;
        double r = s[c][0];
        for(int w = 1; w < 4; w++)
            r = (c < 3) ? ((r < s[c][w]) ? s[c][w] : r) : ((s[c][w] < r) ? s[c][w] : r);
        scene[c] = r;
    }
}

__device__ __forceinline__ uint32_t expand_bits(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ double clamp_ref(double v, double lo, double hi)
{
    double a = (v < lo) ? lo : v;  // std::max(v, lo)
    return (hi < a) ? hi : a;      // std::min(a, hi)
}

__device__ __forceinline__ uint32_t morton_code(double x, double y, double z)
{
    const double res = 1024.0;
    x                = clamp_ref(x * res, 0.0, res - 1.0);
    y                = clamp_ref(y * res, 0.0, res - 1.0);
    z                = clamp_ref(z * res, 0.0, res - 1.0);
    return (expand_bits((uint32_t)x) << 2) + (expand_bits((uint32_t)y) << 1)
           + expand_bits((uint32_t)z);
}

__global__ void k_morton(int n, const double* __restrict__ leaf, const double* __restrict__ scene,
                         uint64_t* __restrict__ keys)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if(e >= n)
        return;
    double lo[3] = {scene[3], scene[4], scene[5]};
    double sz[3] = {scene[0] - lo[0], scene[1] - lo[1], scene[2] - lo[2]};
    double c[3];
    for(int k = 0; k < 3; k++)
        c[k] = (leaf[6 * (size_t)e + k] + leaf[6 * (size_t)e + 3 + k]) * 0.5;
    uint64_t mc = morton_code((c[0] - lo[0]) / sz[0], (c[1] - lo[1]) / sz[1], (c[2] - lo[2]) / sz[2]);
    keys[e]     = (mc << 32) | (uint64_t)(uint32_t)e;
}

// ------------------------------------------------------------ radix sort
__global__ void __launch_bounds__(kSortThreads) k_sort_hist(int n, int shift,
                                                            const uint64_t* __restrict__ keys,
                                                            unsigned* __restrict__ hist, int nblk)
{
    __shared__ unsigned h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int base = blockIdx.x * kSortTile;
    for(int j = 0; j < kSortItems; j++)
    {
        int i = base + j * kSortThreads + threadIdx.x;
        if(i < n)
            atomicAdd(&h[(keys[i] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of hist (digit-major, 256 * nblk entries: a multiple of 4),
// one workgroup: tiles of 4096 entries, 4 consecutive per thread (one 16-byte
// load, coalesced across the workgroup), the next tile's load issued before
// the current tile is scanned; a running carry joins the tiles
__global__ void __launch_bounds__(1024) k_sort_scan(int total, unsigned* __restrict__ hist)
{
    __shared__ unsigned wsum[16];
    uint4*     h4    = reinterpret_cast<uint4*>(hist);
    const int  n4    = total >> 2;
    const int  t     = threadIdx.x;
    const int  w     = t >> 6;
    unsigned   carry = 0;
    uint4      cur   = t < n4 ? h4[t] : make_uint4(0u, 0u, 0u, 0u);
    for(int b4 = 0; b4 < n4; b4 += 1024)
    {
        const int i4  = b4 + t;
        uint4     nxt = (i4 + 1024 < n4) ? h4[i4 + 1024] : make_uint4(0u, 0u, 0u, 0u);
        unsigned  s0 = cur.x, s1 = s0 + cur.y, s2 = s1 + cur.z, s3 = s2 + cur.w;
        unsigned  s  = s3;  // inclusive scan of the per-thread sums
        for(int o = 1; o < 64; o <<= 1)
        {
            unsigned u = __shfl_up(s, o, 64);
            if((int)__lane_id() >= o)
                s += u;
        }
        if(__lane_id() == 63)
            wsum[w] = s;
        __syncthreads();
        if(t < 16)
        {
            unsigned ws = wsum[t];
            for(int o = 1; o < 16; o <<= 1)
            {
                unsigned u = __shfl_up(ws, o, 16);
                if(t >= o)
                    ws += u;
            }
            wsum[t] = ws;
        }
        __syncthreads();
        unsigned excl = s - s3 + (w > 0 ? wsum[w - 1] : 0u) + carry;
        if(i4 < n4)
            h4[i4] = make_uint4(excl, excl + s0, excl + s1, excl + s2);
        carry += wsum[15];
        __syncthreads();  // wsum is rewritten by the next tile
        cur = nxt;
    }
}

// stable scatter: keys are visited in tile order (round j, then thread);
// with KV each key's 32-bit value moves with it
template <bool KV>
__global__ void __launch_bounds__(kSortThreads) k_sort_scatter(int n, int shift,
                                                               const uint64_t* __restrict__ in,
                                                               uint64_t* __restrict__ out,
                                                               const unsigned* __restrict__ hist,
                                                               int nblk,
                                                               const uint32_t* __restrict__ vin,
                                                               uint32_t* __restrict__ vout)
{
    __shared__ unsigned base_d[256];
    __shared__ unsigned wcnt[4][256];
    const int lane = __lane_id();
    const int w    = threadIdx.x >> 6;
    base_d[threadIdx.x] = hist[(size_t)threadIdx.x * nblk + blockIdx.x];
    const int tile      = blockIdx.x * kSortTile;
    for(int j = 0; j < kSortItems; j++)
    {
        for(int q = 0; q < 4; q++)
            wcnt[q][threadIdx.x] = 0;
        __syncthreads();
        int      i    = tile + j * kSortThreads + threadIdx.x;
        bool     live = i < n;
        uint64_t k    = live ? in[i] : 0;
        unsigned d    = (unsigned)((k >> shift) & 0xFF);
        unsigned long long peers = __ballot(live);
#pragma unroll
        for(int b = 0; b < 8; b++)
        {
            unsigned long long m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        unsigned long long lt   = (lane == 0) ? 0ull : (peers & ((~0ull) >> (64 - lane)));
        unsigned           rank = (unsigned)__popcll(lt);
        if(live && rank == 0)
            wcnt[w][d] = (unsigned)__popcll(peers);
        __syncthreads();
        if(live)
        {
            unsigned off = base_d[d] + rank;
            for(int q = 0; q < w; q++)
                off += wcnt[q][d];
            out[off] = k;
            if(KV)
                vout[off] = vin[i];
        }
        __syncthreads();
        base_d[threadIdx.x] += wcnt[0][threadIdx.x] + wcnt[1][threadIdx.x] + wcnt[2][threadIdx.x]
                               + wcnt[3][threadIdx.x];
        __syncthreads();
    }
}

// --------------------------------------------------------- Karras build
__device__ __forceinline__ int cub64(uint64_t a, uint64_t b) { return __clzll(a ^ b); }

__global__ void k_leaf_nodes(int n, const uint64_t* __restrict__ keys, uint32_t* __restrict__ nodes,
                             const double* __restrict__ leaf, double* __restrict__ bvs,
                             unsigned* __restrict__ flags)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n)
        return;
    if(i < n - 1)
    {
        uint32_t* nd = nodes + 4 * (size_t)i;
        nd[0] = nd[1] = nd[2] = nd[3] = 0xFFFFFFFFu;
        flags[i]                      = 0u;
    }
    uint32_t  el = (uint32_t)(keys[i] & 0xFFFFFFFFull);
    uint32_t* nd = nodes + 4 * (size_t)(i + n - 1);
    nd[0] = nd[1] = nd[2] = 0xFFFFFFFFu;
    nd[3]                 = el;
    for(int c = 0; c < 6; c++)
        bvs[6 * (size_t)(i + n - 1) + c] = leaf[6 * (size_t)el + c];
}

__global__ void k_internal_nodes(int n, const uint64_t* __restrict__ code, uint32_t* __restrict__ nodes,
                                 int* __restrict__ rng)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if(idx >= n - 1)
        return;
    int first, last;
    if(idx == 0)
    {
        first = 0;
        last  = n - 1;
    }
    else
    {
        uint64_t self = code[idx];
        int      Ld   = cub64(self, code[idx - 1]);
        int      Rd   = cub64(self, code[idx + 1]);
        int      d    = (Rd > Ld) ? 1 : -1;
        int      dmin = Ld < Rd ? Ld : Rd;
        int      lmax = 2;
        int      delta = -1;
        int      it    = idx + d * lmax;
        if(0 <= it && it < n)
            delta = cub64(self, code[it]);
        while(delta > dmin)
        {
            lmax <<= 1;
            it    = idx + d * lmax;
            delta = -1;
            if(0 <= it && it < n)
                delta = cub64(self, code[it]);
        }
        int l = 0;
        for(int t = lmax >> 1; t > 0; t >>= 1)
        {
            it    = idx + (l + t) * d;
            delta = -1;
            if(0 <= it && it < n)
                delta = cub64(self, code[it]);
            if(delta > dmin)
                l += t;
        }
        int j = idx + l * d;
        first = d < 0 ? j : idx;
        last  = d < 0 ? idx : j;
    }
    int gamma;
    {
        uint64_t fc = code[first], lc = code[last];
        if(fc == lc)
            gamma = (first + last) >> 1;
        else
        {
            int dn     = cub64(fc, lc);
            int split  = first;
            int stride = last - first;
            do
            {
                stride     = (stride + 1) >> 1;
                int middle = split + stride;
                if(middle < last && cub64(fc, code[middle]) > dn)
                    split = middle;
            } while(stride > 1);
            gamma = split;
        }
    }
    rng[2 * (size_t)idx]     = first < last ? first : last;
    rng[2 * (size_t)idx + 1] = first < last ? last : first;
    uint32_t L = (uint32_t)gamma, R = (uint32_t)gamma + 1;
    if(min(first, last) == gamma)
        L += n - 1;
    if(max(first, last) == gamma + 1)
        R += n - 1;
    nodes[4 * (size_t)idx + 1] = L;
    nodes[4 * (size_t)idx + 2] = R;
    nodes[4 * (size_t)L]       = (uint32_t)idx;
    nodes[4 * (size_t)R]       = (uint32_t)idx;
}

// ------------------------------------------------------------------ refit
// Internal node i covers the sorted leaves [first_i, last_i] (Karras), so its
// box is the union of those leaf boxes.  Unions are exact (max/min), so the
// boxes are bit-identical to a bottom-up refit whatever the order.  Instead
// of walking up with atomics and device-wide fences, every node is computed
// independently from leaf boxes and two levels of precomputed aggregates:
// blocks of 64 leaves and super-blocks of 64 blocks.
constexpr int kAgg = 64;

struct Box6
{
    double u[3], l[3];
};

__device__ __forceinline__ Box6 box_empty()
{
    Box6 b;
    for(int c = 0; c < 3; c++)
    {
        b.u[c] = -INFINITY;
        b.l[c] = INFINITY;
    }
    return b;
}

__device__ __forceinline__ void box_add(Box6& b, const double* p)  // p: {upper, lower}
{
    for(int c = 0; c < 3; c++)
    {
        double a = p[c], d = p[3 + c];
        b.u[c]   = (b.u[c] < a) ? a : b.u[c];
        b.l[c]   = (d < b.l[c]) ? d : b.l[c];
    }
}

// one wave per group of 64 input boxes: out[g] = union of in[64 g .. 64 g + 63]
__global__ void __launch_bounds__(256) k_box_groups(int n, const double* __restrict__ in,
                                                    double* __restrict__ out)
{
    int    i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[6];
    if(i < n)
        for(int c = 0; c < 6; c++)
            v[c] = in[6 * (size_t)i + c];
    else
        for(int c = 0; c < 3; c++)
        {
            v[c]     = -INFINITY;
            v[3 + c] = INFINITY;
        }
#pragma unroll
    for(int o = 32; o > 0; o >>= 1)
#pragma unroll
        for(int c = 0; c < 3; c++)
        {
            double a = __shfl_xor(v[c], o, 64), d = __shfl_xor(v[3 + c], o, 64);
            v[c]     = (v[c] < a) ? a : v[c];
            v[3 + c] = (d < v[3 + c]) ? d : v[3 + c];
        }
    int g = i >> 6;
    if((i & 63) < 6 && g < (n + 63) / 64)
        out[6 * (size_t)g + (i & 63)] = v[i & 63];
}

// union of in[a .. b] (inclusive)
__device__ __forceinline__ void box_scan(Box6& acc, const double* __restrict__ in, int a, int b)
{
    for(int j = a; j <= b; j++)
        box_add(acc, in + 6 * (size_t)j);
}

__global__ void __launch_bounds__(256) k_refit_query(int n, const int* __restrict__ rng,
                                                     const double* __restrict__ blk,
                                                     const double* __restrict__ sup,
                                                     double* __restrict__ bvs)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n - 1)
        return;
    const double* leaf = bvs + 6 * (size_t)(n - 1);
    int           f = rng[2 * (size_t)i], l = rng[2 * (size_t)i + 1];
    Box6          acc = box_empty();
    if(l - f < 2 * kAgg)
        box_scan(acc, leaf, f, l);
    else
    {
        int fb = (f + kAgg - 1) / kAgg, lb = (l + 1) / kAgg - 1;  // full blocks fb..lb
        box_scan(acc, leaf, f, fb * kAgg - 1);
        box_scan(acc, leaf, (lb + 1) * kAgg, l);
        if(lb - fb < 2 * kAgg)
            box_scan(acc, blk, fb, lb);
        else
        {
            int fs = (fb + kAgg - 1) / kAgg, ls = (lb + 1) / kAgg - 1;  // full super-blocks
            box_scan(acc, blk, fb, fs * kAgg - 1);
            box_scan(acc, blk, (ls + 1) * kAgg, lb);
            box_scan(acc, sup, fs, ls);
        }
    }
    double* o = bvs + 6 * (size_t)i;
    for(int c = 0; c < 3; c++)
    {
        o[c]     = acc.u[c];
        o[3 + c] = acc.l[c];
    }
}

void Lbvh::init(int n_, int arity_, const uint32_t* prims_)
{
    n     = n_;
    arity = arity_;
    prims = prims_;
    if(n <= 0)
        return;
    leaf.resize(6 * (size_t)n);
    bvs.resize(6 * (size_t)(2 * n - 1));
    nodes.resize(4 * (size_t)(2 * n - 1));
    keys.resize(n);
    keys_alt.resize(n);
    flags.resize(n);
    rng.resize(2 * (size_t)(n > 1 ? n - 1 : 1));
    int nblk_agg = (n + kAgg - 1) / kAgg;
    agg_blk.resize(6 * (size_t)nblk_agg);
    agg_sup.resize(6 * (size_t)((nblk_agg + kAgg - 1) / kAgg));
    hist.resize(radix_sort_hist_size(n));
    scene.resize(6);
    partial.resize(6 * (size_t)grid_for(n, 256));
}

// exclusive scan of the digit-major histogram (256 digits x nblk tiles), in
// place: one workgroup for short histograms, the three-launch tiled scan
// (tile sums, their scan, apply; its tile sums after the histogram) for long
// ones, where one workgroup's sequential rounds take longer than two launches
static void sort_scan(unsigned* hist, int nblk, hipStream_t s)
{
    const int total = 256 * nblk;
    if(nblk < 128)
        hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(1024), 0, s, total, hist);
    else
        scan_exclusive_u32(total, hist, hist, hist + total, s);
}

// stable LSD radix sort of n 64-bit keys on the 8-bit digits from bit lo to
// bit hi (exclusive); the result ends in keys (alt: scratch of n keys, hist:
// radix_sort_hist_size(n) counters)
void radix_sort_keys(uint64_t* keys, uint64_t* alt, unsigned* hist, int n, int lo, int hi,
                     hipStream_t s)
{
    if(n <= 1)
        return;
    int       nblk = (n + kSortTile - 1) / kSortTile;
    uint64_t* a    = keys;
    uint64_t* b    = alt;
    for(int sh = lo; sh < hi; sh += 8)
    {
        hipLaunchKernelGGL(k_sort_hist, dim3(nblk), dim3(kSortThreads), 0, s, n, sh, a, hist, nblk);
        sort_scan(hist, nblk, s);
        hipLaunchKernelGGL(k_sort_scatter<false>, dim3(nblk), dim3(kSortThreads), 0, s, n, sh, a, b,
                           hist, nblk, nullptr, nullptr);
        SG_LAUNCH_CHECK();
        std::swap(a, b);
    }
    if(a != keys)
        SG_HIP(hipMemcpyAsync(keys, a, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
}

void radix_sort_pairs(uint64_t* keys, uint64_t* kalt, uint32_t* vals, uint32_t* valt,
                      unsigned* hist, int n, int lo, int hi, hipStream_t s)
{
    if(n <= 1)
        return;
    int       nblk = (n + kSortTile - 1) / kSortTile;
    uint64_t* a    = keys;
    uint64_t* b    = kalt;
    uint32_t* va   = vals;
    uint32_t* vb   = valt;
    for(int sh = lo; sh < hi; sh += 8)
    {
        hipLaunchKernelGGL(k_sort_hist, dim3(nblk), dim3(kSortThreads), 0, s, n, sh, a, hist, nblk);
        sort_scan(hist, nblk, s);
        hipLaunchKernelGGL(k_sort_scatter<true>, dim3(nblk), dim3(kSortThreads), 0, s, n, sh, a, b,
                           hist, nblk, va, vb);
        SG_LAUNCH_CHECK();
        std::swap(a, b);
        std::swap(va, vb);
    }
    if(a != keys)
    {
        SG_HIP(hipMemcpyAsync(keys, a, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
        SG_HIP(hipMemcpyAsync(vals, va, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
}

int radix_key_bits(uint64_t max_key)
{
    int b = 0;
    while(b < 64 && (max_key >> b))
        b++;
    return (b + 7) / 8 * 8;
}

size_t radix_sort_hist_size(int n)
{
    const int total = 256 * ((n + kSortTile - 1) / kSortTile);
    return (size_t)total + scan_tmp_size(total);
}

void Lbvh::build(const double* x, const double* dx, double alpha, hipStream_t s)
{
    if(n <= 0)
        return;
    hipLaunchKernelGGL(k_leaf_aabb, dim3(grid_for(n, 256)), dim3(256), 0, s, n, arity, prims, x,
                       dx, alpha, leaf.data());
    unsigned nb = grid_for(n, 256);
    hipLaunchKernelGGL(k_box_partial, dim3(nb), dim3(256), 0, s, n, leaf.data(), partial.data());
    hipLaunchKernelGGL(k_box_final, dim3(1), dim3(256), 0, s, (int)nb, partial.data(), scene.data());
    hipLaunchKernelGGL(k_morton, dim3(grid_for(n, 256)), dim3(256), 0, s, n, leaf.data(),
                       scene.data(), keys.data());
    SG_LAUNCH_CHECK();
    // the keys enter in index order and every pass is stable, so sorting the
    // 30 Morton bits alone orders by (morton, index)
    radix_sort_keys(keys.data(), keys_alt.data(), hist.data(), n, 32, 62, s);
    hipLaunchKernelGGL(k_leaf_nodes, dim3(grid_for(n, 256)), dim3(256), 0, s, n, keys.data(),
                       nodes.data(), leaf.data(), bvs.data(), flags.data());
    if(n > 1)
    {
        hipLaunchKernelGGL(k_internal_nodes, dim3(grid_for(n - 1, 256)), dim3(256), 0, s, n,
                           keys.data(), nodes.data(), rng.data());
        int nblk_agg = (n + kAgg - 1) / kAgg;
        hipLaunchKernelGGL(k_box_groups, dim3(grid_for(n, 256)), dim3(256), 0, s, n,
                           bvs.data() + 6 * (size_t)(n - 1), agg_blk.data());
        hipLaunchKernelGGL(k_box_groups, dim3(grid_for(nblk_agg, 256)), dim3(256), 0, s,
                           nblk_agg, agg_blk.data(), agg_sup.data());
        hipLaunchKernelGGL(k_refit_query, dim3(grid_for(n - 1, 256)), dim3(256), 0, s, n,
                           rng.data(), agg_blk.data(), agg_sup.data(), bvs.data());
    }
    SG_LAUNCH_CHECK();
}

}  // namespace sgipc
