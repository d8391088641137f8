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

    // blk: row-major block (a, b) of the element matrix for vertices ra, rb
    __device__ void put(int ra, int rb, double blk[9], bool diag) const
    {
        bool zero = btype[ra] != 0 || btype[rb] != 0;
        if(zero)
            for(int q = 0; q < 9; q++)
                blk[q] = 0.0;
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
        unsigned o = atomicAdd(coo_count, 2u);
        if((int)(o + 2) > coo_cap)
        {
            atomicExch(overflow, 1);
            return;
        }
        coo_r[o]     = ra;
        coo_c[o]     = rb;
        coo_r[o + 1] = rb;
        coo_c[o + 1] = ra;
        for(int r = 0; r < 3; r++)
            for(int c = 0; c < 3; c++)
            {
                coo_v[9 * (size_t)o + r * 3 + c]       = blk[r * 3 + c];
                coo_v[9 * (size_t)(o + 1) + c * 3 + r] = blk[r * 3 + c];
            }
    }
};

inline BlockSink block_sink(const SimDev& d)
{
    return BlockSink{d.btype, d.rowptr, d.cols, d.vals, d.coo_r, d.coo_c, d.coo_v,
                     d.coo_count, d.coo_cap, d.overflow};
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
