/*
 * orc_lbvh.c -- linear BVH restatement.  TEST INFRASTRUCTURE ONLY.
 *
 * Follows StiffGIPC/mlbvh.cu:
 *   expand_bits / morton_code     :52-77   (10-bit interleave, resolution 1024)
 *   _calcLeafBvs                  :1149-1172
 *   calcMaxBV (scene box)         :1745-1775
 *   _calcMChash                   :1208-1227 (key = morton30 << 32 | index)
 *   thrust::sort_by_key           :2082
 *   _calcLeafNodes                :1229-1247 (leaves at [n-1, 2n-1))
 *   _calcInternalNodes            :1249-1269 (determine_range :135-196,
 *                                             find_split :198-229)
 *   _calcInternalAABB             :1271-1298 (bottom-up refit)
 * Node layout {parent, left, right, element}; AABB layout {upper, lower}.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

uint32_t orc_expand_bits(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

static double clampd(double v, double lo, double hi)
{
    /* std::min(std::max(v, lo), hi) */
    double a = (v < lo) ? lo : v;
    return (hi < a) ? hi : a;
}

uint32_t orc_morton_code(double x, double y, double z)
{
    const double res = 1024.0;
    x                = clampd(x * res, 0.0, res - 1.0);
    y                = clampd(y * res, 0.0, res - 1.0);
    z                = clampd(z * res, 0.0, res - 1.0);
    uint32_t xx      = orc_expand_bits((uint32_t)x);
    uint32_t yy      = orc_expand_bits((uint32_t)y);
    uint32_t zz      = orc_expand_bits((uint32_t)z);
    return (xx << 2) + (yy << 1) + zz;
}

void orc_leaf_aabbs(const double* x, const uint32_t* prims, int arity, int n, double* aabb)
{
    for(int e = 0; e < n; e++)
    {
        double up[3] = {-1e32, -1e32, -1e32}, lo[3] = {1e32, 1e32, 1e32};
        for(int k = 0; k < arity; k++)
        {
            const double* p = x + 3 * prims[arity * e + k];
            for(int c = 0; c < 3; c++)
            {
                lo[c] = (p[c] < lo[c]) ? p[c] : lo[c];
                up[c] = (up[c] < p[c]) ? p[c] : up[c];
            }
        }
        for(int c = 0; c < 3; c++)
        {
            aabb[6 * e + c]     = up[c];
            aabb[6 * e + 3 + c] = lo[c];
        }
    }
}

static int cmp_u64(const void* a, const void* b)
{
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}

static int cub(uint64_t a, uint64_t b) { return __builtin_clzll(a ^ b); }

static void determine_range(const uint64_t* code, int n, int idx, int* first, int* last)
{
    if(idx == 0)
    {
        *first = 0;
        *last  = n - 1;
        return;
    }
    uint64_t self   = code[idx];
    int      Ld     = cub(self, code[idx - 1]);
    int      Rd     = cub(self, code[idx + 1]);
    int      d      = (Rd > Ld) ? 1 : -1;
    int      dmin   = Ld < Rd ? Ld : Rd;
    int      lmax   = 2;
    int      delta  = -1;
    int      it     = idx + d * lmax;
    if(0 <= it && it < n)
        delta = cub(self, code[it]);
    while(delta > dmin)
    {
        lmax <<= 1;
        it    = idx + d * lmax;
        delta = -1;
        if(0 <= it && it < n)
            delta = cub(self, code[it]);
    }
    int l = 0;
    for(int t = lmax >> 1; t > 0; t >>= 1)
    {
        it    = idx + (l + t) * d;
        delta = -1;
        if(0 <= it && it < n)
            delta = cub(self, code[it]);
        if(delta > dmin)
            l += t;
    }
    int j = idx + l * d;
    if(d < 0)
    {
        *first = j;
        *last  = idx;
    }
    else
    {
        *first = idx;
        *last  = j;
    }
}

static int find_split(const uint64_t* code, int first, int last)
{
    uint64_t fc = code[first], lc = code[last];
    if(fc == lc)
        return (first + last) >> 1;
    int dnode  = cub(fc, lc);
    int split  = first;
    int stride = last - first;
    do
    {
        stride     = (stride + 1) >> 1;
        int middle = split + stride;
        if(middle < last)
        {
            if(cub(fc, code[middle]) > dnode)
                split = middle;
        }
    } while(stride > 1);
    return split;
}

void orc_lbvh_build(const double* x, const uint32_t* prims, int arity, int n,
                    uint64_t* keys, uint32_t* nodes, double* aabbs)
{
    double* leaf = (double*)malloc(sizeof(double) * 6 * n);
    orc_leaf_aabbs(x, prims, arity, n, leaf);
    orc_lbvh_build_boxes(leaf, n, keys, nodes, aabbs);
    free(leaf);
}

void orc_lbvh_build_boxes(const double* leaf, int n, uint64_t* keys, uint32_t* nodes,
                          double* aabbs)
{
    int nn = 2 * n - 1;
    double up[3] = {-1e32, -1e32, -1e32}, lo[3] = {1e32, 1e32, 1e32};
    for(int e = 0; e < n; e++)
        for(int c = 0; c < 3; c++)
        {
            up[c] = (up[c] < leaf[6 * e + c]) ? leaf[6 * e + c] : up[c];
            lo[c] = (leaf[6 * e + 3 + c] < lo[c]) ? leaf[6 * e + 3 + c] : lo[c];
        }
    double size[3] = {up[0] - lo[0], up[1] - lo[1], up[2] - lo[2]};
    for(int e = 0; e < n; e++)
    {
        double c[3];
        for(int k = 0; k < 3; k++)
            c[k] = (leaf[6 * e + k] + leaf[6 * e + 3 + k]) * 0.5;
        uint64_t mc = orc_morton_code((c[0] - lo[0]) / size[0], (c[1] - lo[1]) / size[1],
                                      (c[2] - lo[2]) / size[2]);
        keys[e] = (mc << 32) | (uint64_t)(uint32_t)e;
    }
    qsort(keys, n, sizeof(uint64_t), cmp_u64);

    for(int i = 0; i < nn; i++)
    {
        nodes[4 * i + 0] = 0xFFFFFFFFu;
        nodes[4 * i + 1] = 0xFFFFFFFFu;
        nodes[4 * i + 2] = 0xFFFFFFFFu;
        nodes[4 * i + 3] = 0xFFFFFFFFu;
    }
    for(int i = 0; i < n; i++)
    {
        uint32_t el              = (uint32_t)(keys[i] & 0xFFFFFFFFu);
        nodes[4 * (i + n - 1) + 3] = el;
        memcpy(aabbs + 6 * (i + n - 1), leaf + 6 * el, sizeof(double) * 6);
    }
    for(int i = 0; i < n - 1; i++)
    {
        int f, l;
        determine_range(keys, n, i, &f, &l);
        int      g = find_split(keys, f, l);
        uint32_t L = (uint32_t)g, R = (uint32_t)g + 1;
        if((f < l ? f : l) == g)
            L += n - 1;
        if((f > l ? f : l) == g + 1)
            R += n - 1;
        nodes[4 * i + 1]     = L;
        nodes[4 * i + 2]     = R;
        nodes[4 * L + 0]     = (uint32_t)i;
        nodes[4 * R + 0]     = (uint32_t)i;
    }
    /* bottom-up refit: the second child to arrive merges */
    int* visits = (int*)calloc(nn, sizeof(int));
    for(int i = 0; i < n; i++)
    {
        uint32_t p = nodes[4 * (i + n - 1) + 0];
        while(p != 0xFFFFFFFFu)
        {
            if(visits[p]++ == 0)
                break;
            uint32_t L = nodes[4 * p + 1], R = nodes[4 * p + 2];
            for(int c = 0; c < 3; c++)
            {
                double a = aabbs[6 * L + c], b = aabbs[6 * R + c];
                aabbs[6 * p + c] = (a < b) ? b : a;
                a                = aabbs[6 * L + 3 + c];
                b                = aabbs[6 * R + 3 + c];
                aabbs[6 * p + 3 + c] = (b < a) ? b : a;
            }
            p = nodes[4 * p + 0];
        }
    }
    if(n == 1)
        memcpy(aabbs, leaf, sizeof(double) * 6);
    free(visits);
}
