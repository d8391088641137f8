// assembly.h -- device helpers shared by the contact assembly kernels:
// block accumulation, contact-code decoding, static-pattern lookup with a COO
// fallback, and fixed-order block partial sums.
#pragma once

#include "dmath.h"
#include "sim_kernels.h"

namespace sgipc
{

__device__ __forceinline__ void add_block(double* dst, const double b[9], bool transpose)
{
    if(!transpose)
    {
#pragma unroll
        for(int e = 0; e < 9; e++)
            atomicAdd(dst + e, b[e]);
    }
    else
    {
#pragma unroll
        for(int r = 0; r < 3; r++)
#pragma unroll
            for(int c = 0; c < 3; c++)
                atomicAdd(dst + 3 * r + c, b[3 * c + r]);
    }
}

// decode a contact code into type (0 PP, 1 PE, 2 PT, 3 EE) and vertex ids
__device__ __forceinline__ int decode_pair(int4 c, int ids[4])
{
    if(c.x >= 0)
    {
        ids[0] = c.x, ids[1] = c.y, ids[2] = c.z, ids[3] = c.w;
        return 3;
    }
    ids[0] = -c.x - 1;
    ids[1] = c.y;
    if(c.z < 0)
        return 0;
    ids[2] = c.z;
    if(c.w < 0)
        return 1;
    ids[3] = c.w;
    return 2;
}

__device__ __forceinline__ int find_slot(const int* __restrict__ rowptr, const int* __restrict__ cols,
                                         int r, int c)
{
    int lo = rowptr[r], hi = rowptr[r + 1] - 1;
    while(lo <= hi)
    {
        int m  = (lo + hi) >> 1;
        int cm = cols[m];
        if(cm == c)
            return m;
        if(cm < c)
            lo = m + 1;
        else
            hi = m - 1;
    }
    return -1;
}

// Sink for the 3x3 blocks of one contact element: a block (ra, rb) lands in
// the static pattern when (ra, rb) is there (and its transpose in (rb, ra)),
// otherwise in the COO tail as two entries (both orientations).  Blocks
// touching a boundary vertex are zeroed; a zeroed block outside the pattern
// is still appended so the MAS hierarchy sees the contact connectivity.
// Vertices of affine bodies map to their body's 4 nodes: block H between
// vertices i and j becomes C_m H C_n^T between node m of i's body and node n
// of j's (C_0 = I, C_(1+k) = xbar e_k^T), i.e. the blocks of J_i^T H J_j.
struct BlockSink
{
    const int* __restrict__ btype;
    const int* __restrict__ rowptr;
    const int* __restrict__ cols;
    double* __restrict__ vals;
    int* __restrict__ coo_r;
    int* __restrict__ coo_c;
    double* __restrict__ coo_v;
    unsigned* __restrict__ coo_count;
    int  coo_cap;
    int* __restrict__ overflow;
    int  nfem;
    const int* __restrict__ node_of;
    const double* __restrict__ xbar;

    // node block (ra, rb) and, unless diag, its transpose at (rb, ra)
    __device__ void put_node(int ra, int rb, const double blk[9], bool diag, bool zero) const
    {
        int k = find_slot(rowptr, cols, ra, rb);
        if(k >= 0)
        {
            if(!zero)
            {
                add_block(vals + 9 * (size_t)k, blk, false);
                if(!diag)
                    add_block(vals + 9 * (size_t)find_slot(rowptr, cols, rb, ra), blk, true);
            }
            return;
        }
        unsigned o = atomicAdd(coo_count, diag ? 1u : 2u);
        if((int)(o + (diag ? 1 : 2)) > coo_cap)
        {
            atomicExch(overflow, 1);
            return;
        }
        coo_r[o] = ra;
        coo_c[o] = rb;
        for(int q = 0; q < 9; q++)
            coo_v[9 * (size_t)o + q] = zero ? 0.0 : blk[q];
        if(diag)
            return;
        coo_r[o + 1] = rb;
        coo_c[o + 1] = ra;
        for(int r = 0; r < 3; r++)
            for(int c = 0; c < 3; c++)
                coo_v[9 * (size_t)(o + 1) + c * 3 + r] = zero ? 0.0 : blk[r * 3 + c];
    }

    // blk: row-major block (a, b) of the element matrix for vertices ra, rb
    __device__ void put(int ra, int rb, double blk[9], bool diag) const
    {
        bool zero = btype[ra] != 0 || btype[rb] != 0;
        if(ra < nfem && rb < nfem)
        {
            put_node(ra, rb, blk, diag, zero);
            return;
        }
        int na = node_of[ra], nb = node_of[rb];
        int ma = ra < nfem ? 1 : 4, mb = rb < nfem ? 1 : 4;
        double xa[3] = {0, 0, 0}, xb[3] = {0, 0, 0};
        if(ra >= nfem)
            xa[0] = xbar[3 * ra], xa[1] = xbar[3 * ra + 1], xa[2] = xbar[3 * ra + 2];
        if(rb >= nfem)
            xb[0] = xbar[3 * rb], xb[1] = xbar[3 * rb + 1], xb[2] = xbar[3 * rb + 2];
        for(int m = 0; m < ma; m++)
            for(int n = 0; n < mb; n++)
            {
                // C_m blk C_n^T: C_0 = I, C_(1+k) rows = xbar_l e_k^T
                double t[9], o[9];
                for(int i = 0; i < 3; i++)
                    for(int c = 0; c < 3; c++)
                        t[i * 3 + c] = (m == 0) ? blk[i * 3 + c] : xa[i] * blk[(m - 1) * 3 + c];
                for(int i = 0; i < 3; i++)
                    for(int j = 0; j < 3; j++)
                        o[i * 3 + j] = (n == 0) ? t[i * 3 + j] : t[i * 3 + (n - 1)] * xb[j];
                put_node(na + m, nb + n, o, diag, zero);
            }
    }
};

inline BlockSink block_sink(const SimDev& d)
{
    return BlockSink{d.btype, d.rowptr, d.cols, d.vals, d.coo_r, d.coo_c, d.coo_v,
                     d.coo_count, d.coo_cap, d.overflow, d.nfem, d.node_of, d.xbar};
}

// block partial sums (256 threads) written in block order
__device__ __forceinline__ void block_partial(double v, double* out)
{
    __shared__ double ws[4];
    v = wave_sum(v);
    if(__lane_id() == 0)
        ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if(threadIdx.x == 0)
        out[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

}  // namespace sgipc
