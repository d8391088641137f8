/*
 * orc_sort.c -- spatial re-sorting keys.  TEST INFRASTRUCTURE ONLY.
 *
 * Follows StiffGIPC/GIPC.cu:
 *   expand_bits / hash_code   :142-198 (type -1: 30-bit Morton; types 0..5:
 *                                       10-bit cells packed in an axis order)
 *   _calcVertMChash           :412-461 (vertex position, type -1)
 *   _calcTetMChash            :279-341 (centroid 0.25 ((a + b) + (c + d)),
 *                                       type from the box's extent order)
 *   key = code << 32 | index, sorted ascending (thrust::sort_by_key on the
 *   unique keys).  The box is the vertices' AABB.  A 0/0 coordinate (an
 *   axis of zero extent) is cell 0 here; the reference leaves it to an
 *   undefined float-to-integer cast.
 */
#include <stdlib.h>
#include "oracle.h"

static double cell(double v)
{
    const double res = 1024.0;
    v                = v * res;
    double a         = (v > 0.0) ? v : 0.0;
    return (res - 1.0 < a) ? res - 1.0 : a;
}

uint32_t orc_hash_code(int type, double x, double y, double z)
{
    uint32_t X = (uint32_t)cell(x), Y = (uint32_t)cell(y), Z = (uint32_t)cell(z);
    switch(type)
    {
        case -1: return (orc_expand_bits(X) << 2) + (orc_expand_bits(Y) << 1) + orc_expand_bits(Z);
        case 0: return (Z * 1024 + Y) * 1024 + X;
        case 1: return (Y * 1024 + Z) * 1024 + X;
        case 2: return (X * 1024 + Z) * 1024 + Y;
        case 3: return (Z * 1024 + X) * 1024 + Y;
        case 4: return (Y * 1024 + X) * 1024 + Z;
        default: return (X * 1024 + Y) * 1024 + Z;
    }
}

int orc_tet_hash_type(double sx, double sy, double sz)
{
    if(sx > sy && sy > sz)
        return 0;
    if(sx > sz && sz > sy)
        return 1;
    if(sy > sz && sz > sx)
        return 2;
    if(sy > sx && sx > sz)
        return 3;
    if(sz > sx && sx > sy)
        return 4;
    return 5;
}

static int cmp_key(const void* a, const void* b)
{
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}

int orc_spatial_keys(int kind, const double* x, int nv, const uint32_t* tets, int nt,
                     uint64_t* keys, uint32_t* order)
{
    if(nv <= 0 || (kind == 1 && nt <= 0))
        return -1;
    double up[3] = {-1e32, -1e32, -1e32}, lo[3] = {1e32, 1e32, 1e32};
    for(int i = 0; i < nv; i++)
        for(int c = 0; c < 3; c++)
        {
            double p = x[3 * (size_t)i + c];
            up[c]    = (up[c] < p) ? p : up[c];
            lo[c]    = (p < lo[c]) ? p : lo[c];
        }
    double s[3] = {up[0] - lo[0], up[1] - lo[1], up[2] - lo[2]};
    int    n    = kind == 0 ? nv : nt;
    for(int i = 0; i < n; i++)
    {
        uint32_t code;
        if(kind == 0)
        {
            const double* p = x + 3 * (size_t)i;
            code            = orc_hash_code(-1, (p[0] - lo[0]) / s[0], (p[1] - lo[1]) / s[1],
                                            (p[2] - lo[2]) / s[2]);
        }
        else
        {
            const uint32_t* t = tets + 4 * (size_t)i;
            double          c[3];
            for(int k = 0; k < 3; k++)
                c[k] = ((x[3 * (size_t)t[0] + k] + x[3 * (size_t)t[1] + k])
                        + (x[3 * (size_t)t[2] + k] + x[3 * (size_t)t[3] + k]))
                       * 0.25;
            code = orc_hash_code(orc_tet_hash_type(s[0], s[1], s[2]), (c[0] - lo[0]) / s[0],
                                 (c[1] - lo[1]) / s[1], (c[2] - lo[2]) / s[2]);
        }
        keys[i] = ((uint64_t)code << 32) | (uint32_t)i;
    }
    qsort(keys, (size_t)n, sizeof(uint64_t), cmp_key);
    if(order)
        for(int i = 0; i < n; i++)
            order[i] = (uint32_t)(keys[i] & 0xffffffffu);
    return 0;
}
