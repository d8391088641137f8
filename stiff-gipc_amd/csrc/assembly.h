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
// otherwise in the contact table as two entries (both orientations), merged
// with every other block of the same (row, col) on insertion (affine-body
// contacts put many blocks on the same node pairs).  Blocks touching a
// boundary vertex are zeroed; a zeroed block outside the pattern still
// creates its entry so the MAS hierarchy sees the contact connectivity.
// Vertices of affine bodies map to their body's 4 nodes: block H between
// vertices i and j becomes C_m H C_n^T between node m of i's body and node n
// of j's (C_0 = I, C_(1+k) = xbar e_k^T), i.e. the blocks of J_i^T H J_j.
struct BlockSink
{
    const int* __restrict__ btype;
    const int* __restrict__ rowptr;
    const int* __restrict__ cols;
    double* __restrict__ vals;
    unsigned long long* __restrict__ hkeys;
    double* __restrict__ hvals;
    int* __restrict__ hlist;
    unsigned* __restrict__ hcount;
    int* __restrict__ hfail;
    unsigned long long hmask;
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
        table_add(ra, rb, blk, false, zero);
        if(!diag)
            table_add(rb, ra, blk, true, zero);
    }

    // add blk (or its transpose) into the contact table entry (r, c): linear
    // probing; the thread that claims an empty slot records it in the slot
    // list; a full table sets hfail (the assembly is then redone with a
    // larger table)
    __device__ void table_add(int r, int c, const double blk[9], bool transposed,
                              bool zero) const
    {
        unsigned long long key = ((unsigned long long)(unsigned)r << 32) | (unsigned)c;
        unsigned long long h   = mix64(key) & hmask;
        for(unsigned long long probe = 0; probe <= hmask; probe++)
        {
            unsigned long long prev = atomicCAS(hkeys + h, kEmptyKey, key);
            if(prev == kEmptyKey)
                hlist[atomicAdd(hcount, 1u)] = (int)h;
            if(prev == kEmptyKey || prev == key)
            {
                if(!zero)
                    for(int i = 0; i < 3; i++)
                        for(int j = 0; j < 3; j++)
                        {
                            double x = transposed ? blk[j * 3 + i] : blk[i * 3 + j];
                            if(x != 0.0)
                                atomicAdd(hvals + 9 * h + i * 3 + j, x);
                        }
                return;
            }
            h = (h + 1) & hmask;
        }
        atomicExch(hfail, 1);
    }

    // Element matrix of the form H_ab = c_a c_b K (a, b < nv; K a symmetric
    // 3x3), which every contact Hessian here is (barrier: c_a = w_a / sqrt(dhat),
    // K = lambda0 u u^T; friction: c_a = w_a, K = the tangent block).  The
    // element's vertices are grouped by node set first (an FEM vertex is its
    // own node; the vertices of one affine body share the body's 4 nodes), and
    // J^T H J is formed per pair of groups: with W = sum c_a and X = sum c_a xbar_a
    // over a group, node block (m, n) is C_m K C_n^T for C_0 = W I and
    // C_(1+k) = X e_k^T.  For FEM vertices this is the per-vertex block; for
    // affine bodies it sums the blocks of all the element's vertices in a body
    // before they reach the contact table (one entry per node pair instead of
    // one per vertex pair).
    __device__ void put_rank(const int* ids, int nv, const double* c, const double K[9]) const
    {
        int    gbase[4], gn[4], ng = 0;
        double gw[4], gx[4][3];
        bool   gz[4];
        for(int a = 0; a < nv; a++)
        {
            int  v   = ids[a];
            bool abd = v >= nfem;
            int  nb  = abd ? node_of[v] : v;
            int  g   = 0;
            while(g < ng && gbase[g] != nb)
                g++;
            if(g == ng)
            {
                gbase[g] = nb;
                gn[g]    = abd ? 4 : 1;
                gw[g]    = 0.0;
                gx[g][0] = gx[g][1] = gx[g][2] = 0.0;
                gz[g]    = false;
                ng++;
            }
            gw[g] += c[a];
            if(abd)
                for(int k = 0; k < 3; k++)
                    gx[g][k] += c[a] * xbar[3 * (size_t)v + k];
            gz[g] = gz[g] || btype[v] != 0;
        }
        for(int g1 = 0; g1 < ng; g1++)
            for(int g2 = g1; g2 < ng; g2++)
            {
                bool zero = gz[g1] || gz[g2];
                for(int m = 0; m < gn[g1]; m++)
                    for(int n = 0; n < gn[g2]; n++)
                    {
                        double o[9];
                        for(int i = 0; i < 3; i++)
                            for(int j = 0; j < 3; j++)
                            {
                                // (C_m K C_n^T)[i][j]
                                double left = (m == 0) ? 0.0 : gx[g1][i];
                                double v;
                                if(m == 0 && n == 0)
                                    v = gw[g1] * gw[g2] * K[i * 3 + j];
                                else if(m == 0)
                                    v = gw[g1] * K[i * 3 + (n - 1)] * gx[g2][j];
                                else if(n == 0)
                                    v = left * K[(m - 1) * 3 + j] * gw[g2];
                                else
                                    v = left * K[(m - 1) * 3 + (n - 1)] * gx[g2][j];
                                o[i * 3 + j] = v;
                            }
                        // within one group every (m, n) is written once
                        // (diag); across groups the transpose goes to (n, m)
                        put_node(gbase[g1] + m, gbase[g2] + n, o, g1 == g2, zero);
                    }
            }
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
    return BlockSink{d.btype, d.rowptr, d.cols, d.vals, d.hkeys, d.hvals, d.hlist, d.hcount,
                     d.hfail, d.hmask, d.overflow, d.nfem, d.node_of, d.xbar};
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
