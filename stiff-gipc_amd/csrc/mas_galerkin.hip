// mas_galerkin.hip -- the MAS bank matrices of the coarse levels, summed in a
// fixed order.
//
// The bank matrix of level l >= 1 holds the blocks (ci, cj) of the Galerkin
// operator A_l = P_l^T A P_l whose two nodes share a bank (ci >> 4 == cj >> 4;
// levels are 16-aligned in the global node numbering), where P_l aggregates
// every node into its level-l cluster (PrepareHessian_bcoo,
// MASPreconditioner.cu:1810-2084, sums the same blocks).  Level by level:
// A_l = the blocks of A_(l-1) keyed by (parent(row), parent(col)), sorted by
// key (stable radix sort: equal keys stay in the order of the entries of
// A_(l-1)) and each key's run summed in that order.  A_0 is the stored
// matrix: the static BSR in row-major order, then the contact blocks in the
// order the contact table lists them; the sums do not depend on timing
// beyond that order.
#include "linsys.h"
#include "sim_kernels.h"

namespace sgipc
{

namespace
{
constexpr int kScanTile = 1024;  // 256 threads x 4 entries

__global__ void k_gal_keys0(int nnz, int coo_cap, const unsigned* __restrict__ coo_count,
                            const int* __restrict__ rows, const int* __restrict__ cols,
                            const int* __restrict__ cr, const int* __restrict__ cc,
                            const int* __restrict__ chain, int Lm1, int off1, long long n1,
                            uint64_t kdead, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= nnz + coo_cap)
        return;
    int ncoo = coo_count ? (int)*coo_count : 0;
    ncoo     = ncoo < coo_cap ? ncoo : coo_cap;
    int i = -1, j = -1;
    if(k < nnz)
        i = rows[k], j = cols[k];
    else if(k - nnz < ncoo)
        i = cr[k - nnz], j = cc[k - nnz];
    uint64_t key = kdead;
    if(i >= 0)
        key = (uint64_t)(chain[(size_t)i * Lm1] - off1) * (uint64_t)n1
              + (uint64_t)(chain[(size_t)j * Lm1] - off1);
    keys[k] = key;
    vals[k] = (uint32_t)k;
}

__global__ void k_gal_keys(int n, const int* __restrict__ er, const int* __restrict__ ec,
                           const int* __restrict__ going_next, int offl, long long nl,
                           uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= n)
        return;
    keys[k] = (uint64_t)(going_next[er[k]] - offl) * (uint64_t)nl
              + (uint64_t)(going_next[ec[k]] - offl);
    vals[k] = (uint32_t)k;
}

__global__ void k_gal_heads(int n, const uint64_t* __restrict__ keys, uint64_t kdead,
                            unsigned* __restrict__ flag)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= n)
        return;
    const uint64_t key = keys[k];
    flag[k]            = (key != kdead && (k == 0 || keys[k - 1] != key)) ? 1u : 0u;
}

// run starts in run order: rstart[pos[k]] = k for every run head k
__global__ void k_gal_starts(int n, const unsigned* __restrict__ flag,
                             const unsigned* __restrict__ pos, int* __restrict__ rstart)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k < n && flag[k])
        rstart[pos[k]] = k;
}

// exclusive scan of n unsigned values: tile sums, a one-workgroup scan of the
// tile sums (total in tsum[nt]), then each tile rescanned with its offset
__device__ __forceinline__ unsigned block_scan_excl(unsigned v, unsigned* ws, unsigned& total)
{
    const int lane = __lane_id(), w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    unsigned  s    = v;
    for(int o = 1; o < 64; o <<= 1)
    {
        unsigned u = __shfl_up(s, o, 64);
        if(lane >= o)
            s += u;
    }
    if(lane == 63)
        ws[w] = s;
    __syncthreads();
    unsigned before = 0;
    total           = 0;
    for(int i = 0; i < nw; i++)
    {
        if(i < w)
            before += ws[i];
        total += ws[i];
    }
    __syncthreads();
    return before + s - v;
}

__global__ void __launch_bounds__(256) k_scan_tiles(int n, const unsigned* __restrict__ in,
                                                    unsigned* __restrict__ tsum)
{
    __shared__ unsigned ws[4];
    const int base = blockIdx.x * kScanTile + 4 * threadIdx.x;
    unsigned  v    = 0;
    for(int q = 0; q < 4; q++)
        v += (base + q < n) ? in[base + q] : 0u;
    unsigned total;
    (void)block_scan_excl(v, ws, total);
    if(threadIdx.x == 0)
        tsum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(1024) k_scan_top(int nt, unsigned* __restrict__ tsum)
{
    __shared__ unsigned ws[16];
    unsigned            carry = 0;
    for(int b = 0; b < nt; b += 1024)
    {
        const int i = b + threadIdx.x;
        unsigned  v = i < nt ? tsum[i] : 0u;
        unsigned  total;
        unsigned  e = block_scan_excl(v, ws, total);
        if(i < nt)
            tsum[i] = carry + e;
        carry += total;
    }
    if(threadIdx.x == 0)
        tsum[nt] = carry;
}

// in and out may be the same array: each thread reads its entries before the
// scan and writes them after it
__global__ void __launch_bounds__(256) k_scan_apply(int n, const unsigned* in,
                                                    const unsigned* __restrict__ tsum,
                                                    unsigned* out)
{
    __shared__ unsigned ws[4];
    const int base = blockIdx.x * kScanTile + 4 * threadIdx.x;
    unsigned  v[4], t = 0;
    for(int q = 0; q < 4; q++)
    {
        v[q] = (base + q < n) ? in[base + q] : 0u;
        t += v[q];
    }
    unsigned total;
    unsigned e = block_scan_excl(t, ws, total) + tsum[blockIdx.x];
    for(int q = 0; q < 4; q++)
    {
        if(base + q < n)
            out[base + q] = e;
        e += v[q];
    }
}

// one thread per run (runs packed by their start, the count in *nruns): the
// run's blocks in sorted order (ties in entry order), four entries' indices
// and blocks requested at a time, stored as entry r of A_l and, within a
// bank, into B
__global__ void k_gal_reduce(int n, const unsigned* __restrict__ nruns,
                             const int* __restrict__ rstart, const uint64_t* __restrict__ keys,
                             const uint32_t* __restrict__ vals, int nsrc0,
                             const double* __restrict__ src0, const double* __restrict__ src1,
                             int offl, long long nl, int* __restrict__ er, int* __restrict__ ec,
                             double* __restrict__ eb, double* __restrict__ B)
{
    const int      r   = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned cnt = *nruns;
    if(r >= (int)cnt)
        return;
    const int      k   = rstart[r];
    const uint64_t key = keys[k];
    int            e   = (r + 1 < (int)cnt) ? rstart[r + 1] : n;
    if(r + 1 == (int)cnt)  // the last run ends where the dead keys start
    {
        int m = k + 1;
        while(m < n && keys[m] == key)
            m++;
        e = m;
    }
    double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    auto   src  = [&](uint32_t v) {
        return (int)v < nsrc0 ? src0 + 9 * (size_t)v : src1 + 9 * (size_t)(v - nsrc0);
    };
    int m = k;
    for(; m + 4 <= e; m += 4)
    {
        const double* blk[4];
        double        val[4][9];
#pragma unroll
        for(int q = 0; q < 4; q++)
            blk[q] = src(vals[m + q]);
#pragma unroll
        for(int q = 0; q < 4; q++)
#pragma unroll
            for(int c = 0; c < 9; c++)
                val[q][c] = blk[q][c];
#pragma unroll
        for(int q = 0; q < 4; q++)
#pragma unroll
            for(int c = 0; c < 9; c++)
                s[c] += val[q][c];
    }
    for(; m < e; m++)
    {
        const double* blk = src(vals[m]);
#pragma unroll
        for(int c = 0; c < 9; c++)
            s[c] += blk[c];
    }
    const int ci = offl + (int)(key / (uint64_t)nl);
    const int cj = offl + (int)(key % (uint64_t)nl);
    er[r]        = ci;
    ec[r]        = cj;
#pragma unroll
    for(int c = 0; c < 9; c++)
        eb[9 * (size_t)r + c] = s[c];
    if((ci >> 4) == (cj >> 4))
    {
        double* M = B + (size_t)(ci >> 4) * 48 * 48 + (3 * (ci & 15)) * 48 + 3 * (cj & 15);
#pragma unroll
        for(int a = 0; a < 3; a++)
#pragma unroll
            for(int c = 0; c < 3; c++)
                M[a * 48 + c] = s[3 * a + c];
    }
}
}  // namespace

void scan_exclusive_u32(int n, const unsigned* in, unsigned* out, unsigned* tsum, hipStream_t s)
{
    if(n <= 0)
        return;
    const int nt = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(256), 0, s, n, in, tsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, nt, tsum);
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(256), 0, s, n, in, tsum, out);
    SG_LAUNCH_CHECK();
}

size_t scan_tmp_size(int n) { return (size_t)(n + kScanTile - 1) / kScanTile + 1; }

void MasGalerkin::ensure(int n)
{
    if(n <= cap_)
        return;
    cap_ = n;
    keys_.resize(n);
    kalt_.resize(n);
    vals_.resize(n);
    valt_.resize(n);
    flag_.resize(n);
    pos_.resize(n);
    rstart_.resize(n);
    hist_.resize(radix_sort_hist_size(n));
    tsum_.resize((n + kScanTile - 1) / kScanTile + 1);
    for(int q = 0; q < 2; q++)
    {
        er_[q].resize(n);
        ec_[q].resize(n);
        eb_[q].resize(9 * (size_t)n);
    }
}

// sort the n keys (with their entry indices), find the runs, number them and
// sum them into entry buffer `out`; returns the number of entries
int MasGalerkin::reduce(int n, uint64_t kdead, int nsrc0, const double* src0, const double* src1,
                        int offl, long long nl, double* B, int out, hipStream_t s)
{
    if(n <= 0)
        return 0;
    radix_sort_pairs(keys_.data(), kalt_.data(), vals_.data(), valt_.data(), hist_.data(), n, 0,
                     radix_key_bits(kdead), s);
    const unsigned g  = grid_for(n, 256);
    const int      nt = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_gal_heads, dim3(g), dim3(256), 0, s, n, keys_.data(), kdead,
                       flag_.data());
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(256), 0, s, n, flag_.data(), tsum_.data());
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, nt, tsum_.data());
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(256), 0, s, n, flag_.data(), tsum_.data(),
                       pos_.data());
    hipLaunchKernelGGL(k_gal_starts, dim3(g), dim3(256), 0, s, n, flag_.data(), pos_.data(),
                       rstart_.data());
    hipLaunchKernelGGL(k_gal_reduce, dim3(g), dim3(256), 0, s, n, tsum_.data() + nt,
                       rstart_.data(), keys_.data(), vals_.data(), nsrc0, src0, src1, offl, nl,
                       er_[out].data(), ec_[out].data(), eb_[out].data(), B);
    SG_LAUNCH_CHECK();
    unsigned cnt = 0;
    SG_HIP(hipMemcpyAsync(&cnt, tsum_.data() + nt, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    return (int)cnt;
}

void MasGalerkin::run(const Bsr& A, const Coo* coo, int L, const std::vector<int>& lvl,
                      const int* chain, const int* going_next, double* B, hipStream_t s)
{
    if(L < 2)
        return;
    const int coo_cap = coo ? (coo->live ? coo->live : coo->cap) : 0;
    int       n       = A.nnzb + coo_cap;
    ensure(n);
    long long nl   = lvl[2 * 1];
    int       offl = lvl[2 * 1 + 1];
    uint64_t  kd   = (uint64_t)nl * (uint64_t)nl;
    hipLaunchKernelGGL(k_gal_keys0, dim3(grid_for(n, 256)), dim3(256), 0, s, A.nnzb, coo_cap,
                       coo ? coo->count.data() : nullptr, A.rows.data(), A.cols.data(),
                       coo ? coo->r.data() : nullptr, coo ? coo->c.data() : nullptr, chain, L - 1,
                       offl, nl, kd, keys_.data(), vals_.data());
    SG_LAUNCH_CHECK();
    int cur = 0;
    n       = reduce(n, kd, A.nnzb, A.vals.data(), coo ? coo->v.data() : nullptr, offl, nl, B, cur,
                     s);
    for(int l = 2; l < L; l++)
    {
        nl   = lvl[2 * l];
        offl = lvl[2 * l + 1];
        kd   = (uint64_t)nl * (uint64_t)nl;
        hipLaunchKernelGGL(k_gal_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, n,
                           er_[cur].data(), ec_[cur].data(), going_next, offl, nl, keys_.data(),
                           vals_.data());
        SG_LAUNCH_CHECK();
        n   = reduce(n, kd, n, eb_[cur].data(), nullptr, offl, nl, B, 1 - cur, s);
        cur = 1 - cur;
    }
}

}  // namespace sgipc
