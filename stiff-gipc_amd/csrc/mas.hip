// mas.hip -- multilevel additive Schwarz (MAS) preconditioner on gfx950.
//
// Semantics follow StiffGIPC/MASPreconditioner.cu and main.cu:47-75:
//   * bank p holds the vertices of part p in slots 16p .. 16p+15 (run order);
//   * levelnum = computeNumLevels(nv) (MASPreconditioner.cu:1720-1739);
//   * the nodes of level l+1 are the connected components of the graph induced
//     on each 16-node bank of level l (mesh adjacency plus contact cliques),
//     numbered bank by bank in order of each component's lowest lane
//     (_buildCML0_new :63, _preparePrefixSumL0_new :157, _buildLevel1_new :271,
//     _buildConnectMaskLx_new :401, _nextLevelCluster :480, _prefixSumLx :530);
//   * goingNext[g] = global id of node g one level up (_buildLevel1_new,
//     _prefixSumLx), chain[v] = global ids of vertex v at levels 1..L-1
//     (_aggregationKernel :592);
//   * bank matrices = (P_l^T A P_l) restricted to each bank, zero diagonals set
//     to 1, inverted (PrepareHessian_bcoo :1810-2084, __inverse6_P96x96 :628);
//   * z = sum_l P_l Binv_l P_l^T r (__buildMultiLevelR_optimized_new :727,
//     _schwarzLocalXSym6 :954, __collectFinalZ_new :848).
//
// MI355X design: a 16-node bank is a quarter of a wave64; closures run on a
// bank's 16 lanes through LDS, counts come from one 64-bit ballot.  Coarse
// Galerkin sums use wave-aggregated atomics (one atomic per distinct target
// per wave).  Each bank matrix is inverted by one wave in LDS (lane = column).
// The apply is level-synchronous and atomic-free: every node of level l+1 lies
// inside a single bank of level l, so restriction is a bank-local LDS
// reduction done by the same wave that applies that bank's inverse.
#include "linsys.h"
#include "sgipc.h"

namespace sgipc
{

constexpr int BANK = 16;
constexpr int BD   = 48;  // 3 * BANK
constexpr int ZB   = 64;  // Z entries per bank: one per lane, lanes 48..63 padding

static int compute_num_levels(int nv)
{
    int nLevel  = 1;
    int levelSz = (nv + BANK - 1) / BANK * BANK;
    while(levelSz > BANK)
    {
        levelSz /= BANK;
        nLevel++;
        levelSz = (levelSz + BANK - 1) / BANK * BANK;
    }
    nLevel = nLevel + 1;
    return nLevel > 6 ? 6 : nLevel;
}

// ------------------------------------------------------------ hierarchy
__global__ void k_l0_mask_coo(const unsigned* __restrict__ count, unsigned cap,
                              const int* __restrict__ cr, const int* __restrict__ cc,
                              const int* __restrict__ slot_of, unsigned* mask)
{
    unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned n = *count;
    if(t >= (n < cap ? n : cap))
        return;
    int a = slot_of[cr[t]], b = slot_of[cc[t]];
    if(a != b && (a >> 4) == (b >> 4))
        atomicOr(mask + a, 1u << (b & 15));
}

// mask[key] |= bits for the lanes with key >= 0, merging equal keys of a
// wave (up to 8 distinct keys; further lanes use their own atomic)
__device__ __forceinline__ void wave_or_keyed(unsigned* mask, int key, unsigned bits)
{
    unsigned long long todo = __ballot(key >= 0);
    for(int round = 0; round < 8 && todo; round++)
    {
        int      leader = __ffsll((long long)todo) - 1;
        int      k      = __shfl(key, leader, 64);
        bool     mine   = key == k;
        unsigned v      = mine ? bits : 0u;
#pragma unroll
        for(int o = 32; o > 0; o >>= 1)
            v |= __shfl_xor(v, o, 64);
        if((int)__lane_id() == leader)
            atomicOr(mask + k, v);
        todo &= ~__ballot(mine);
    }
    if(todo & (1ull << __lane_id()))
        atomicOr(mask + key, bits);
}

// level l >= 1: mask bits from static cross-bank edges and COO entries
__global__ void k_lx_mask(int nx, const int* __restrict__ xa, const int* __restrict__ xb,
                          const unsigned* __restrict__ coo_count, int coo_cap,
                          const int* __restrict__ cr, const int* __restrict__ cc,
                          const int* __restrict__ node_l, unsigned* mask)
{
    int t     = blockIdx.x * blockDim.x + threadIdx.x;
    int ncoo  = coo_count ? (int)*coo_count : 0;
    ncoo      = ncoo < coo_cap ? ncoo : coo_cap;
    int a = -1, b = -1;  // every lane stays active for the wave merge below
    if(t < nx + ncoo)
    {
        int u = t < nx ? xa[t] : cr[t - nx];
        int v = t < nx ? xb[t] : cc[t - nx];
        a     = node_l[u];
        b     = node_l[v];
        if(!(a != b && (a >> 4) == (b >> 4)))
            a = b = -1;
    }
    // At coarse levels most edges of a wave land on the same few nodes: the
    // bits of equal targets are OR-ed across the wave first, then one atomic
    // per target
    wave_or_keyed(mask, a, 1u << (b & 15));
    wave_or_keyed(mask, b, 1u << (a & 15));
}

__device__ __forceinline__ int level_count(const int* lvl, int level, int total_map)
{
    return level == 0 ? total_map : lvl[2 * level];
}

// closure of the 16-lane bank masks; per-bank cluster counts
__global__ void __launch_bounds__(256) k_cluster(int level, int total_map,
                                                 const int* __restrict__ lvl,
                                                 const int* __restrict__ vert_of_slot,
                                                 const unsigned* __restrict__ mask,
                                                 unsigned* __restrict__ comp,
                                                 int* __restrict__ bank_cnt)
{
    __shared__ unsigned sm[256];
    int  t      = blockIdx.x * blockDim.x + threadIdx.x;
    int  lane16 = t & 15;
    int  n      = level_count(lvl, level, total_map);
    int  nbanks = (n + 15) >> 4;
    bool inb    = (t >> 4) < nbanks;
    bool valid  = inb && (level == 0 ? vert_of_slot[t] >= 0 : t < n);
    unsigned m  = valid ? (mask[t] | (1u << lane16)) : 0u;
    sm[threadIdx.x] = m;
    __syncthreads();
    unsigned c = m, visited = valid ? (1u << lane16) : 0u;
    int      gb = threadIdx.x & ~15;
    while(true)
    {
        unsigned todo = c & ~visited;
        if(!todo)
            break;
        int nx = __ffs(todo) - 1;
        visited |= 1u << nx;
        c |= sm[gb + nx];
    }
    bool               elected = valid && ((c & ((1u << lane16) - 1u)) == 0u);
    unsigned long long bal     = __ballot(elected);
    unsigned           grp     = (unsigned)((bal >> (threadIdx.x & 48)) & 0xFFFFull);
    if(inb)
    {
        comp[t] = c;
        if(lane16 == 0)
            bank_cnt[t >> 4] = __popc(grp);
    }
}

// exclusive scan of bank counts (one workgroup); writes level l+1 size/offset
__global__ void __launch_bounds__(1024) k_scan_banks(int level, int total_map, int* lvl,
                                                     const int* __restrict__ bank_cnt,
                                                     int* __restrict__ bank_prefix)
{
    __shared__ int wsum[16];
    __shared__ int carry_s;
    int n      = level_count(lvl, level, total_map);
    int nbanks = (n + 15) >> 4;
    if(threadIdx.x == 0)
        carry_s = 0;
    __syncthreads();
    for(int base = 0; base < nbanks; base += 1024)
    {
        int i = base + threadIdx.x;
        int v = (i < nbanks) ? bank_cnt[i] : 0;
        // inclusive wave scan
        int s = v;
        for(int o = 1; o < 64; o <<= 1)
        {
            int u = __shfl_up(s, o, 64);
            if((int)__lane_id() >= o)
                s += u;
        }
        int w = threadIdx.x >> 6;
        if(__lane_id() == 63)
            wsum[w] = s;
        __syncthreads();
        if(threadIdx.x < 16)
        {
            int ws = wsum[threadIdx.x];
            for(int o = 1; o < 16; o <<= 1)
            {
                int u = __shfl_up(ws, o, 16);
                if((int)threadIdx.x >= o)
                    ws += u;
            }
            wsum[threadIdx.x] = ws;
        }
        __syncthreads();
        int carry = carry_s;
        int excl  = s - v + (w > 0 ? wsum[w - 1] : 0) + carry;
        if(i < nbanks)
            bank_prefix[i] = excl;
        __syncthreads();
        if(threadIdx.x == 1023)
            carry_s = excl + v;
        __syncthreads();
    }
    if(threadIdx.x == 0)
    {
        int total            = carry_s;
        int off              = lvl[2 * level + 1];
        lvl[2 * (level + 1)]     = total;
        lvl[2 * (level + 1) + 1] = off + ((n + 15) / 16) * 16;
    }
}

// new (level l+1) local ids, goingNext and per-bank first/count
__global__ void __launch_bounds__(256) k_assign(int level, int total_map, const int* __restrict__ lvl,
                                                const int* __restrict__ vert_of_slot,
                                                const unsigned* __restrict__ comp,
                                                const int* __restrict__ bank_prefix,
                                                const int* __restrict__ bank_cnt,
                                                int* __restrict__ newid, int* __restrict__ going_next,
                                                int* __restrict__ bank_first, int* __restrict__ bank_ncl,
                                                int* __restrict__ child_bank)
{
    int  t      = blockIdx.x * blockDim.x + threadIdx.x;
    int  lane16 = t & 15;
    int  n      = level_count(lvl, level, total_map);
    int  nbanks = (n + 15) >> 4;
    bool inb    = (t >> 4) < nbanks;
    bool valid  = inb && (level == 0 ? vert_of_slot[t] >= 0 : t < n);
    unsigned c  = inb ? comp[t] : 0u;
    bool     elected = valid && ((c & ((1u << lane16) - 1u)) == 0u);
    unsigned long long bal = __ballot(elected);
    unsigned grp           = (unsigned)((bal >> (threadIdx.x & 48)) & 0xFFFFull);
    if(!inb)
        return;
    int off_l  = lvl[2 * level + 1];
    int off_n  = lvl[2 * (level + 1) + 1];
    int bank   = t >> 4;
    if(lane16 == 0)
    {
        bank_first[(off_l >> 4) + bank] = off_n + bank_prefix[bank];
        bank_ncl[(off_l >> 4) + bank]   = bank_cnt[bank];
    }
    if(!valid)
    {
        newid[t] = -1;
        return;
    }
    int e    = __ffs(c) - 1;
    int id   = bank_prefix[bank] + __popc(grp & ((1u << e) - 1u));
    newid[t] = id;
    going_next[off_l + t] = off_n + id;
    // the bank holding the members of the new node (every member stores the
    // same value)
    child_bank[off_n + id] = (off_l >> 4) + bank;
}

// per fine vertex: node id at level l+1 and its chain entry
__global__ void k_chain(int level, int L, int nv, const int* __restrict__ lvl,
                        const int* __restrict__ slot_of, const int* __restrict__ newid,
                        int* __restrict__ node, int* __restrict__ chain)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if(v >= nv)
        return;
    int src = (level == 0) ? slot_of[v] : node[(size_t)level * nv + v];
    int nid = newid[src];
    node[(size_t)(level + 1) * nv + v]   = nid;
    chain[(size_t)v * (L - 1) + level] = lvl[2 * (level + 1) + 1] + nid;
}

// ------------------------------------------------------------- assembly
// Level 0: every stored block (i, j) of A (both triangles, plus the COO tail)
// whose nodes share a level-0 bank is that bank's (i, j) block; every
// (row, col) block appears once (static pattern and merged contact blocks
// are disjoint), so each bank entry has a single writer.  The coarse levels'
// bank matrices come from MasGalerkin (mas_galerkin.hip).
__global__ void __launch_bounds__(256) k_mas_gather(int nnz, const int* __restrict__ rows,
                                                    const int* __restrict__ cols,
                                                    const double* __restrict__ vals,
                                                    const unsigned* __restrict__ coo_count,
                                                    int coo_cap, const int* __restrict__ cr,
                                                    const int* __restrict__ cc,
                                                    const double* __restrict__ cv,
                                                    const int* __restrict__ slot_of,
                                                    double* __restrict__ B)
{
    int  t    = blockIdx.x * blockDim.x + threadIdx.x;
    int  ncoo = coo_count ? (int)*coo_count : 0;
    ncoo      = ncoo < coo_cap ? ncoo : coo_cap;
    int           i = 0, j = 0;
    const double* blk = nullptr;
    if(t < nnz)
    {
        i   = rows[t];
        j   = cols[t];
        blk = vals + 9 * (size_t)t;
    }
    else if(t < nnz + ncoo)
    {
        i   = cr[t - nnz];
        j   = cc[t - nnz];
        blk = cv + 9 * (size_t)(t - nnz);
    }
    else
        return;
    int si = slot_of[i], sj = slot_of[j];
    if((si >> 4) != (sj >> 4))
        return;
    double* M = B + (size_t)(si >> 4) * BD * BD + (3 * (si & 15)) * BD + 3 * (sj & 15);
#pragma unroll
    for(int a = 0; a < 3; a++)
#pragma unroll
        for(int c = 0; c < 3; c++)
            M[a * BD + c] = blk[3 * a + c];
}

// Ordering of LDS between the lanes of a single-wave workgroup: the LDS
// instructions of one wave execute in program order, so a compiler-only
// wave barrier suffices; __syncthreads would also make the wave wait for its
// outstanding global stores (workgroup-scope release).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_wave_barrier();
}


// One wave per bank: load 48x48 into LDS, zero diagonal -> 1, in-place
// Gauss-Jordan without pivoting (bank matrices are SPD), blocked by 16x16
// tiles so that three quarters of the work runs on FP64 MFMA.  Eliminating
// the 16 pivots of tile k is, in exact arithmetic, the block step
//   P = A_kk,  A_kj <- P^-1 A_kj,  A_ij <- A_ij - A_ik P^-1 A_kj,
//   A_ik <- -A_ik P^-1,  A_kk <- P^-1          (i, j != k),
// the same operator as the 48 scalar pivot steps.  P^-1 is formed by those
// scalar steps on the tile alone (same per-element operations as an
// unblocked elimination); the products are 16x16x16 MFMA tiles
// (v_mfma_f64_16x16x4_f64, four per tile product).
//
// MFMA f64 16x16x4 operand layout (lane l, g = l / 16, j = l % 16):
// A operand A[j][g], B operand B[g][j], accumulator element v of lane l is
// C[g + 4 v][j] (rows interleaved across the lane groups).
constexpr int kInvLd = BD + 1;  // LDS row stride (doubles) of the bank matrix
typedef double dbl4 __attribute__((ext_vector_type(4)));

// acc += sgn * X Y over one 16-deep tile product, X at (xr, xc), Y at
// (yr, yc) of the LDS matrix S
__device__ __forceinline__ dbl4 tile_mma(const double* S, int xr, int xc, int yr, int yc,
                                         double sgn, dbl4 acc)
{
    const int l = __lane_id();
    const int g = l >> 4, j = l & 15;
#pragma unroll
    for(int q = 0; q < 4; q++)
    {
        double a = sgn * S[(xr + j) * kInvLd + xc + 4 * q + g];
        double b = S[(yr + 4 * q + g) * kInvLd + yc + j];
        acc      = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    return acc;
}

__device__ __forceinline__ dbl4 tile_load(const double* S, int r0, int c0)
{
    const int l = __lane_id();
    const int g = l >> 4, j = l & 15;
    dbl4      t;
#pragma unroll
    for(int v = 0; v < 4; v++)
        t[v] = S[(r0 + g + 4 * v) * kInvLd + c0 + j];
    return t;
}

__device__ __forceinline__ void tile_store(double* S, int r0, int c0, dbl4 t)
{
    const int l = __lane_id();
    const int g = l >> 4, j = l & 15;
#pragma unroll
    for(int v = 0; v < 4; v++)
        S[(r0 + g + 4 * v) * kInvLd + c0 + j] = t[v];
}

template <typename T>
__global__ void __launch_bounds__(64) k_mas_invert(const double* __restrict__ B, T* __restrict__ inv)
{
    __shared__ double S[BD * kInvLd];
    const int     bank = blockIdx.x;
    const int     l    = threadIdx.x;
    const int     g    = l >> 4, j = l & 15;
    const double* src  = B + (size_t)bank * BD * BD;
    for(int e = l; e < BD * BD; e += 64)
    {
        int    r = e / BD, c = e % BD;
        double v = src[e];
        S[r * kInvLd + c] = (r == c && v == 0.0) ? 1.0 : v;
    }
    wave_lds_sync();
#pragma unroll
    for(int k = 0; k < 3; k++)
    {
        const int o = 16 * k;
        // P^-1 in registers: lane l holds P[g + 4 v][j]
        dbl4 t = tile_load(S, o, o);
#pragma unroll
        for(int p = 0; p < 16; p++)
        {
            const int pg = p & 3, pv = p >> 2;  // row p = pg + 4 pv
            double    prow = __shfl(t[pv], 16 * pg + j, 64);  // P[p][j]
            double    piv  = __shfl(t[pv], 16 * pg + p, 64);  // P[p][p]
            double    pcol[4];
#pragma unroll
            for(int v = 0; v < 4; v++)
                pcol[v] = __shfl(t[v], 16 * g + p, 64);  // P[g + 4 v][p]
            double ip = 1.0 / piv;
            double a  = prow * ip;
#pragma unroll
            for(int v = 0; v < 4; v++)
            {
                const bool prow_here = (g + 4 * v == p);
                if(j == p)
                    t[v] = prow_here ? ip : -pcol[v] * ip;
                else
                    t[v] = prow_here ? a : t[v] - pcol[v] * a;
            }
        }
        tile_store(S, o, o, t);
        wave_lds_sync();
        // row panel: A_kj <- P^-1 A_kj (each product reads A_kj before its
        // own store; LDS operations of one wave complete in program order)
#pragma unroll
        for(int jt = 0; jt < 3; jt++)
        {
            if(jt == k)
                continue;
            dbl4 u = {0.0, 0.0, 0.0, 0.0};
            u      = tile_mma(S, o, o, o, 16 * jt, 1.0, u);
            wave_lds_sync();
            tile_store(S, o, 16 * jt, u);
            wave_lds_sync();
        }
        // trailing tiles: A_ij -= A_ik (P^-1 A_kj)
#pragma unroll
        for(int it = 0; it < 3; it++)
#pragma unroll
            for(int jt = 0; jt < 3; jt++)
            {
                if(it == k || jt == k)
                    continue;
                dbl4 c = tile_load(S, 16 * it, 16 * jt);
                c      = tile_mma(S, 16 * it, o, o, 16 * jt, -1.0, c);
                wave_lds_sync();
                tile_store(S, 16 * it, 16 * jt, c);
                wave_lds_sync();
            }
        // column panel: A_ik <- -A_ik P^-1
#pragma unroll
        for(int it = 0; it < 3; it++)
        {
            if(it == k)
                continue;
            dbl4 c = {0.0, 0.0, 0.0, 0.0};
            c      = tile_mma(S, 16 * it, o, o, o, -1.0, c);
            wave_lds_sync();
            tile_store(S, 16 * it, o, c);
            wave_lds_sync();
        }
    }
    // Stored transposed (dst[u][t] = Binv[t][u], so that the lanes of a wave
    // read a row with unit stride), keeping only the upper 3x3 blocks of the
    // inverse, as the reference's MasMatrixSymf does (eigen_data.h:125-128,
    // written by __inverse6_P96x96, MASPreconditioner.cu:628-724): below the
    // block diagonal the operator is the transpose of the upper blocks
    T* dst = inv + (size_t)bank * BD * BD;
    for(int e = l; e < BD * BD; e += 64)
    {
        int r = e / BD, c = e % BD;
        dst[e] = (T)((r / 3 >= c / 3) ? S[c * kInvLd + r] : S[r * kInvLd + c]);
    }
}

// ---------------------------------------------------------------- apply
// One bank's product z = Binv R with one wave: lane u < 48 holds R_u, lane
// t < 48 produces z_t (lanes 48..63 compute a copy of lane 0's and store it
// into the bank's padding).  The 48 entries a lane needs (column t of the
// stored transpose) are loaded before the products so their latencies
// overlap; every lane stores unconditionally so the compiler keeps the loads
// together instead of sinking them into a guarded block.
template <typename T>
__device__ __forceinline__ void bank_load(const T* __restrict__ inv, int gb, T (&pv)[BD])
{
    const int t  = __lane_id();
    const int tc = t < BD ? t : 0;
    const T*  P  = inv + (size_t)gb * BD * BD + tc;
#pragma unroll
    for(int u = 0; u < BD; u++)
        pv[u] = P[u * BD];
}

template <typename T>
__device__ __forceinline__ T bank_apply(const T (&pv)[BD], T rv)
{
    T acc = 0;
#pragma unroll
    for(int u = 0; u < BD; u++)
        acc += pv[u] * __shfl(rv, u, 64);
    return acc;
}

template <typename T>
__device__ __forceinline__ T bank_product(const T* __restrict__ inv, int gb, T rv)
{
    T pv[BD];
    bank_load(inv, gb, pv);
    return bank_apply(pv, rv);
}

// Cluster sums of one bank's R into the next level, stored (the bank owns
// those nodes).  rv: lane t < 48 holds component t % 3 of bank node t / 3;
// node_up: global id of that node's parent (or -1).  Lane t < 3 ncl sums
// component t % 3 of the bank's nodes whose parent is cluster first + t / 3,
// in node order, so the sums do not depend on timing.  Single-wave
// workgroups only.
template <typename T>
__device__ __forceinline__ void restrict_bank(T (&Rc)[BD], T rv, int node_up, int first,
                                              int ncl, T* __restrict__ R)
{
    __shared__ int up[BANK];
    const int      t = __lane_id();
    if(t < BD)
        Rc[t] = rv;
    if(t < BD && t % 3 == 0)
        up[t / 3] = node_up;
    wave_lds_sync();
    if(t < 3 * ncl)
    {
        const int p = first + t / 3, c = t % 3;
        T         v = 0;
#pragma unroll
        for(int u = 0; u < BANK; u++)
            if(up[u] == p)
                v += Rc[3 * u + c];
        R[3 * (size_t)first + t] = v;
    }
}

// Level-0 banks (one wave each): optional PCG update of x and r
// (a = rz[k-1] / pAp[k] from slot sets), z0 = Binv r, and the level-1
// cluster sums of r.  T: preconditioner precision, V: precision of the PCG
// vectors.
template <typename T, typename V>
__global__ void __launch_bounds__(64) k_mas_l0(const int* __restrict__ vert_of_slot,
                                               V* __restrict__ x, V* __restrict__ r,
                                               const V* __restrict__ p,
                                               const V* __restrict__ Ap,
                                               const double* __restrict__ rz_prev,
                                               const double* __restrict__ pap,
                                               const T* __restrict__ inv,
                                               const int* __restrict__ going_next,
                                               const int* __restrict__ bank_first,
                                               const int* __restrict__ bank_ncl, T* __restrict__ R,
                                               T* __restrict__ Z, int has_coarse,
                                               const int* __restrict__ done)
{
    if(done && *done)
        return;
    __shared__ T Rc[BD];
    // banks in contiguous ranges per XCD, as the SpMV's block rows: a bank's
    // vertices are then mostly rows the same XCD's SpMV workgroups wrote
    const int b  = (int)xcd_block(blockIdx.x, gridDim.x);
    const int t  = threadIdx.x;
    const int s  = b * BANK + t / 3;
    const int cc = t % 3;
    // issue order: the small loads the vector update depends on, then the
    // bank's inverse, so its fetch overlaps the dependent chain
    // vert_of_slot -> (x, r, p, Ap) -> stores
    int    v  = (t < BD) ? vert_of_slot[s] : -1;
    double ra = 0.0, pa = 0.0;
    if(x)
    {
        ra = rz_prev[t * kSlotStride];
        pa = pap[t * kSlotStride];
    }
    // restriction target: this node's parent, independent of everything
    // before the restriction (the load is unconditional from a valid
    // fallback address, bank_ncl + b, so no branch join forces an early wait)
    int first = 0, ncl = 0;
    if(has_coarse)
    {
        first = bank_first[b];
        ncl   = bank_ncl[b];
    }
    const int gn = *((has_coarse && t < BD) ? going_next + s : bank_ncl + b);
    T         pv[BD];
    bank_load(inv, b, pv);
    double a = 0.0;
    if(x)
        a = wave_sum(ra) / wave_sum(pa);  // slot_set_sum of both sets
    // the updated x and r are stored after the product, so no wait for a
    // store acknowledgement sits between the loads and the product
    V      rv = 0, xn = 0;
    size_t d  = 3 * (size_t)(v >= 0 ? v : 0) + cc;
    if(v >= 0)
    {
        V av = (V)a;
        rv   = r[d];
        if(x)
        {
            xn = x[d] + av * p[d];
            rv -= av * Ap[d];
        }
    }
    // z = Binv r with r broadcast from LDS (12 ds_read_b128 instead of 48
    // cross-lane permutes; same products, same order)
    __shared__ T Rs[64];
    T rt  = (T)rv;
    Rs[t] = rt;
    wave_lds_sync();
    T z = 0;
#pragma unroll
    for(int u = 0; u < BD; u++)
        z += pv[u] * Rs[u];
    Z[(size_t)b * ZB + t] = z;
    if(x && v >= 0)
    {
        x[d] = xn;
        r[d] = rv;
    }
    if(has_coarse)
        restrict_bank(Rc, rt, v >= 0 ? gn : -1, first, ncl, R);
}

// Banks of one coarse level, one wave each: z = Binv R and the next level's
// sums.
template <typename T>
__global__ void __launch_bounds__(64) k_mas_lx(int bank0, int node0, int count,
                                               const T* __restrict__ inv,
                                               const int* __restrict__ going_next,
                                               const int* __restrict__ bank_first,
                                               const int* __restrict__ bank_ncl, T* __restrict__ R,
                                               T* __restrict__ Z, int has_next,
                                               const int* __restrict__ done)
{
    if(done && *done)
        return;
    __shared__ T Rc[BD];
    const int gb    = bank0 + blockIdx.x;
    const int t     = threadIdx.x;
    const int g     = gb * BANK + t / 3;  // global node id
    const int cc    = t % 3;
    const bool valid = (t < BD) && (g - node0 < count);
    T         rv    = valid ? R[3 * (size_t)g + cc] : (T)0;
    T         z     = bank_product(inv, gb, rv);
    Z[(size_t)gb * ZB + t] = z;
    if(has_next)
        restrict_bank(Rc, rv, valid ? going_next[g] : -1, bank_first[gb], bank_ncl[gb], R);
}

// The coarse levels from lt up (the tail), in one launch: every workgroup
// loads R_lt (stored complete by the level lt-1 launch), sums the higher
// levels' R in LDS and gives one wave to each of its share of the tail banks
// (each wave at most one bank: the launch has ceil(banks / kCoarseWaves)
// workgroups, each restricting R redundantly).  Everything it reads from
// global memory -- this wave's bank inverse (FP32), R_lt, goingNext and the
// member banks over the tail's node span -- is requested in one batch per
// thread before the first use, so the chain is one memory round trip, the
// LDS restriction and the product.  Every cluster sum adds the cluster's
// member nodes in node order: the members of cluster p of level l+1 are the
// nodes of the level-l bank cb[p] whose parent is p.
constexpr int kCoarseWaves = 16;
constexpr int kTailBanks   = 128;   // a level with at most this many banks may start the tail
constexpr int kTailNodes   = 4096;  // node span of the tail levels (a multiple of 16)
constexpr int kTailThreads = 64 * kCoarseWaves;
constexpr int kTailSpanPer = kTailNodes / kTailThreads;      // span entries per thread
constexpr int kTailRPer    = 3 * kTailNodes / kTailThreads;  // R values per thread
struct LevelTable
{
    int count[24];
    int off[24];
};

template <typename T>
__global__ void __launch_bounds__(kTailThreads)
    k_mas_coarse(LevelTable lv, int L, int lt, const T* __restrict__ inv,
                 const int* __restrict__ going_next, const int* __restrict__ child_bank,
                 const T* __restrict__ R, T* __restrict__ Z, const int* __restrict__ done)
{
    if(done && *done)
        return;
    __shared__ T             Rs[3 * kTailNodes];
    __shared__ int           up[kTailNodes];
    __shared__ unsigned char cb[kTailNodes];  // member bank of a node above lt, minus b0
    const int w    = threadIdx.x >> 6;
    const int t    = threadIdx.x & 63;
    const int base = lv.off[lt];
    const int span = lv.off[L - 1] + lv.count[L - 1] - base;
    const int nb   = (span + BANK - 1) / BANK;  // banks of the tail levels
    const int b0   = base / BANK;
    const int bi   = blockIdx.x * kCoarseWaves + w;
    const int nr   = 3 * lv.count[lt];          // R_lt values
    const int nz   = 3 * BANK * nb;             // R entries of the tail in LDS
    // FP64 inverses would spill the registers here: loaded after the restriction
    T pv[BD];
    if constexpr(sizeof(T) == 4)
        bank_load(inv, b0 + (bi < nb ? bi : 0), pv);
    const int above = (lt + 1 < L) ? lv.off[lt + 1] - base : span;
    T         rl[kTailRPer];
    int       upv[kTailSpanPer];
    int       cbv[kTailSpanPer];
#pragma unroll
    for(int h = 0; h < kTailRPer; h++)
    {
        const int i = threadIdx.x + h * kTailThreads;
        rl[h]       = i < nr ? R[3 * (size_t)base + i] : (T)0;
    }
#pragma unroll
    for(int h = 0; h < kTailSpanPer; h++)
    {
        const int i = threadIdx.x + h * kTailThreads;
        upv[h]      = i < span ? going_next[base + i] : -1;
        cbv[h]      = (i < span && i >= above) ? child_bank[base + i] - b0 : 0;
    }
#pragma unroll
    for(int h = 0; h < kTailRPer; h++)
    {
        const int i = threadIdx.x + h * kTailThreads;
        if(i < nz)
            Rs[i] = rl[h];
    }
#pragma unroll
    for(int h = 0; h < kTailSpanPer; h++)
    {
        const int i = threadIdx.x + h * kTailThreads;
        if(i < span)
        {
            up[i] = upv[h];
            if(i >= above)
                cb[i] = (unsigned char)cbv[h];
        }
    }
    __syncthreads();
    for(int l = lt; l + 1 < L; l++)
    {
        const int e0 = lv.off[l] + lv.count[l], n1 = lv.off[l + 1];
        for(int q = threadIdx.x; q < lv.count[l + 1]; q += blockDim.x)
        {
            const int p  = n1 + q;
            const int gb = b0 + cb[p - base];
            T         s0 = 0, s1 = 0, s2 = 0;
            // the bank's parents and values read as batches of 8 members
            // (reads past the level's end stay inside the tail's LDS rows;
            // their values are not used), summed in node order
#pragma unroll
            for(int h = 0; h < BANK; h += 8)
            {
                int par[8];
                T   v[24];
#pragma unroll
                for(int u = 0; u < 8; u++)
                {
                    const int g = gb * BANK + h + u;
                    par[u]      = g < e0 ? up[g - base] : -1;
                    v[3 * u]     = Rs[3 * (g - base)];
                    v[3 * u + 1] = Rs[3 * (g - base) + 1];
                    v[3 * u + 2] = Rs[3 * (g - base) + 2];
                }
#pragma unroll
                for(int u = 0; u < 8; u++)
                    if(par[u] == p)
                    {
                        s0 += v[3 * u];
                        s1 += v[3 * u + 1];
                        s2 += v[3 * u + 2];
                    }
            }
            Rs[3 * (p - base)]     = s0;
            Rs[3 * (p - base) + 1] = s1;
            Rs[3 * (p - base) + 2] = s2;
        }
        __syncthreads();
    }
    if(bi < nb)
    {
        const int gb = b0 + bi;
        const T*  rs = Rs + 3 * (gb * BANK - base);
        if constexpr(sizeof(T) == 8)
            bank_load(inv, gb, pv);
        T z = 0;
#pragma unroll
        for(int u = 0; u < BD; u++)
            z += pv[u] * rs[u];
        Z[(size_t)gb * ZB + t] = z;
    }
}

constexpr int kChainFast = 8;  // hierarchies up to 9 levels take an unrolled gather

// z_v = Z[slot(v)] + sum_l Z[chain_l(v)];  r.z into a slot set; workgroup 0
// clears two slot sets for later iterations.  NL = L - 1 chain entries per
// node, known at compile time for NL <= kChainFast (every chain id is
// loaded, then every gather issued: two memory round trips instead of two
// per level; same summation order), or -1 for a runtime count.
template <typename T, typename V, int NL>
__global__ void __launch_bounds__(256)
    k_mas_collect(int nv, int L, const int* __restrict__ slot_of, const int* __restrict__ chain,
                  const T* __restrict__ Z, const V* __restrict__ r, V* __restrict__ z,
                  double* rz_out, double* zero0, double* zero1,
                  const int* __restrict__ done)
{
    if(done && *done)
        return;
    int    v = (int)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    double d = 0.0;
    if(v < nv)
    {
        int        s  = slot_of[v];
        const T*   zs = Z + (size_t)(s >> 4) * ZB + 3 * (s & 15);
        const int* ch = chain + (size_t)v * (L - 1);
        T          z0, z1, z2;
        double     r0 = 0.0, r1 = 0.0, r2 = 0.0;  // loaded with the ids, not after the stores
        if(rz_out)
        {
            r0 = (double)r[3 * (size_t)v];
            r1 = (double)r[3 * (size_t)v + 1];
            r2 = (double)r[3 * (size_t)v + 2];
        }
        if constexpr(NL >= 0)
        {
            int g[NL > 0 ? NL : 1];
#pragma unroll
            for(int l = 0; l < NL; l++)
                g[l] = ch[l];
            z0 = zs[0];
            z1 = zs[1];
            z2 = zs[2];
            T zz[NL > 0 ? NL : 1][3];
#pragma unroll
            for(int l = 0; l < NL; l++)
            {
                const T* zg = Z + (size_t)(g[l] >> 4) * ZB + 3 * (g[l] & 15);
                zz[l][0]    = zg[0];
                zz[l][1]    = zg[1];
                zz[l][2]    = zg[2];
            }
#pragma unroll
            for(int l = 0; l < NL; l++)
            {
                z0 += zz[l][0];
                z1 += zz[l][1];
                z2 += zz[l][2];
            }
        }
        else
        {
            z0 = zs[0];
            z1 = zs[1];
            z2 = zs[2];
            for(int l = 0; l < L - 1; l++)
            {
                const T* zg = Z + (size_t)(ch[l] >> 4) * ZB + 3 * (ch[l] & 15);
                z0 += zg[0];
                z1 += zg[1];
                z2 += zg[2];
            }
        }
        z[3 * (size_t)v]     = (V)z0;
        z[3 * (size_t)v + 1] = (V)z1;
        z[3 * (size_t)v + 2] = (V)z2;
        if(rz_out)
            d = r0 * (double)z0 + r1 * (double)z1 + r2 * (double)z2;
    }
    if(rz_out)
        slot_set_add(rz_out, d);
    if(blockIdx.x == 0)
    {
        if(zero0)
            slot_set_clear(zero0, threadIdx.x);
        if(zero1)
            slot_set_clear(zero1, threadIdx.x);
    }
}

// ----------------------------------------------------------------- host
void Mas::init(const Bsr& A, const int* part_id, const MasConfig& cfg, hipStream_t s)
{
    cfg_ = cfg;
    nv_  = A.nb;
    std::vector<int> slot(nv_);
    int              part = -1, idx = 0, maxpart = -1, prev = -1;
    for(int v = 0; v < nv_; v++)
    {
        int pv = part_id ? part_id[v] : v / BANK;
        SG_CHECK(pv >= prev, -1, "MAS: part ids must be non-decreasing in vertex order");
        prev = pv;
        if(pv != part)
        {
            part = pv;
            idx  = 0;
        }
        SG_CHECK(idx < BANK, -1, "MAS: a part holds more than 16 vertices");
        slot[v] = BANK * pv + idx;
        idx++;
        maxpart = pv > maxpart ? pv : maxpart;
    }
    nparts_    = maxpart + 1;
    total_map_ = BANK * nparts_;
    L_         = compute_num_levels(nv_);
    std::vector<int>      vos(total_map_, -1);
    std::vector<unsigned> m0(total_map_, 0u);
    std::vector<int>      xa, xb;
    for(int v = 0; v < nv_; v++)
        vos[slot[v]] = v;
    for(int i = 0; i < nv_; i++)
        for(int k = A.h_rowptr[i]; k < A.h_rowptr[i + 1]; k++)
        {
            int j = A.h_cols[k];
            if(j == i)
                continue;
            int a = slot[i], b = slot[j];
            if((a >> 4) == (b >> 4))
                m0[a] |= 1u << (b & 15);
            else if(i < j)
            {
                xa.push_back(i);
                xb.push_back(j);
            }
        }
    nx_ = (int)xa.size();
    slot_of_.upload(slot, s);
    vert_of_slot_.upload(vos, s);
    mask0_static_.upload(m0, s);
    xa_.upload(xa, s);
    xb_.upload(xb, s);
    mask_.resize(total_map_);
    comp_.resize(total_map_);
    newid_.resize(total_map_);
    bank_cnt_.resize(total_map_ / BANK + 1);
    bank_prefix_.resize(total_map_ / BANK + 1);
    node_.resize((size_t)L_ * nv_);
    chain_.resize((size_t)nv_ * (L_ > 1 ? L_ - 1 : 1));
    lvl_.resize(2 * (L_ + 2));
    size_t bound = (size_t)L_ * (total_map_ + BANK) + BANK;
    going_next_.resize(bound);
    child_bank_.resize(bound);
    bank_first_.resize(bound / BANK + 1);
    bank_ncl_.resize(bound / BANK + 1);
    std::vector<int> lvl0(2 * (L_ + 2), 0);
    lvl0[0] = nv_;
    lvl_.upload(lvl0, s);
    SG_HIP(hipStreamSynchronize(s));
    built_ = false;
}

void Mas::cluster_level(int level, hipStream_t s)
{
    unsigned g = grid_for(total_map_, 256);
    hipLaunchKernelGGL(k_cluster, dim3(g), dim3(256), 0, s, level, total_map_, lvl_.data(),
                       vert_of_slot_.data(), mask_.data(), comp_.data(), bank_cnt_.data());
    SG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_banks, dim3(1), dim3(1024), 0, s, level, total_map_, lvl_.data(),
                       bank_cnt_.data(), bank_prefix_.data());
    SG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_assign, dim3(g), dim3(256), 0, s, level, total_map_, lvl_.data(),
                       vert_of_slot_.data(), comp_.data(), bank_prefix_.data(), bank_cnt_.data(),
                       newid_.data(), going_next_.data(), bank_first_.data(), bank_ncl_.data(),
                       child_bank_.data());
    SG_LAUNCH_CHECK();
    if(level + 1 <= L_ - 1)
    {
        hipLaunchKernelGGL(k_chain, dim3(grid_for(nv_, 256)), dim3(256), 0, s, level, L_, nv_,
                           lvl_.data(), slot_of_.data(), newid_.data(), node_.data(),
                           chain_.data());
        SG_LAUNCH_CHECK();
    }
}

void Mas::build(const SysView& S, hipStream_t s)
{
    const Coo* coo     = (S.coo && S.coo->cap > 0) ? S.coo : nullptr;
    int        coo_cap = coo ? coo->cap : 0;
    // launches over the contact entries: the live bound when known
    long long  coo_ent = coo ? (coo->live ? (long long)coo->live : (long long)coo_cap) : 0;
    // ---- hierarchy (ReorderRealtime)
    SG_HIP(hipMemcpyAsync(mask_.data(), mask0_static_.data(), total_map_ * sizeof(unsigned),
                          hipMemcpyDeviceToDevice, s));
    if(coo)
    {
        hipLaunchKernelGGL(k_l0_mask_coo, dim3(grid_for(coo_ent, 256)), dim3(256), 0, s,
                           coo->count.data(), (unsigned)coo_cap, coo->r.data(), coo->c.data(),
                           slot_of_.data(), mask_.data());
        SG_LAUNCH_CHECK();
    }
    cluster_level(0, s);
    for(int l = 1; l < L_; l++)
    {
        SG_HIP(hipMemsetAsync(mask_.data(), 0, total_map_ * sizeof(unsigned), s));
        long long ne = (long long)nx_ + coo_ent;
        if(ne > 0)
        {
            hipLaunchKernelGGL(k_lx_mask, dim3(grid_for(ne, 256)), dim3(256), 0, s, nx_,
                               xa_.data(), xb_.data(), coo ? coo->count.data() : nullptr, coo_cap,
                               coo ? coo->r.data() : nullptr, coo ? coo->c.data() : nullptr,
                               node_.data() + (size_t)l * nv_, mask_.data());
            SG_LAUNCH_CHECK();
        }
        cluster_level(l, s);
    }
    h_lvl_.resize(2 * (L_ + 1));
    SG_HIP(hipMemcpyAsync(h_lvl_.data(), lvl_.data(), h_lvl_.size() * sizeof(int),
                          hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    total_clusters_ = h_lvl_[2 * L_ + 1];
    size_t nbanks   = total_clusters_ / BANK;
    if(nbanks > nbanks_alloc_)
    {
        B_.resize(nbanks * BD * BD);
        if(cfg_.fp32)
        {
            inv32_.resize(nbanks * BD * BD);
            R32_.resize(3 * (size_t)total_clusters_);
            Z32_.resize((size_t)ZB * nbanks);
        }
        else
        {
            inv64_.resize(nbanks * BD * BD);
            R64_.resize(3 * (size_t)total_clusters_);
            Z64_.resize((size_t)ZB * nbanks);
        }
        nbanks_alloc_ = nbanks;
    }
    assemble(S, s);
    invert(s);
    // every R entry is stored before it is read; zeroed once all the same
    if(cfg_.fp32)
        R32_.zero(s);
    else
        R64_.zero(s);
    built_ = true;
}

void Mas::assemble(const SysView& S, hipStream_t s)
{
    const Bsr& A       = *S.A;
    const Coo* coo     = (S.coo && S.coo->cap > 0) ? S.coo : nullptr;
    size_t     nbanks  = total_clusters_ / BANK;
    SG_HIP(hipMemsetAsync(B_.data(), 0, nbanks * BD * BD * sizeof(double), s));
    long long n = (long long)A.nnzb + (coo ? (coo->live ? (long long)coo->live : coo->cap) : 0);
    hipLaunchKernelGGL(k_mas_gather, dim3(grid_for(n, 256)), dim3(256), 0, s, A.nnzb,
                       A.rows.data(), A.cols.data(), A.vals.data(),
                       coo ? coo->count.data() : nullptr, coo ? coo->cap : 0,
                       coo ? coo->r.data() : nullptr, coo ? coo->c.data() : nullptr,
                       coo ? coo->v.data() : nullptr, slot_of_.data(), B_.data());
    SG_LAUNCH_CHECK();
    galerkin_.run(A, coo, L_, h_lvl_, chain_.data(), going_next_.data(), B_.data(), s);
}

void Mas::bank_host(int bank, double* mat, double* inv, hipStream_t s) const
{
    SG_CHECK(built_ && bank >= 0 && bank < total_clusters_ / BANK, SGIPC_EINVAL,
             "bank index out of range");
    const size_t off = (size_t)bank * 48 * 48;
    if(mat)
        SG_HIP(hipMemcpyAsync(mat, B_.data() + off, 48 * 48 * sizeof(double),
                              hipMemcpyDeviceToHost, s));
    if(inv && !cfg_.fp32)
        SG_HIP(hipMemcpyAsync(inv, inv64_.data() + off, 48 * 48 * sizeof(double),
                              hipMemcpyDeviceToHost, s));
    std::vector<float> f;
    if(inv && cfg_.fp32)
    {
        f.resize(48 * 48);
        SG_HIP(hipMemcpyAsync(f.data(), inv32_.data() + off, 48 * 48 * sizeof(float),
                              hipMemcpyDeviceToHost, s));
    }
    SG_HIP(hipStreamSynchronize(s));
    for(size_t i = 0; i < f.size(); i++)
        inv[i] = f[i];
}

void Mas::invert(hipStream_t s)
{
    unsigned nbanks = total_clusters_ / BANK;
    if(prof)
        prof->begin("mas_invert", s);
    if(!cfg_.blocked_invert)
    {
        // scalar Gauss-Jordan in the reference's order (mas_invert.hip)
        if(cfg_.fp32)
            launch_mas_invert_gj(B_.data(), inv32_.data(), (int)nbanks, s);
        else
            launch_mas_invert_gj(B_.data(), inv64_.data(), (int)nbanks, s);
    }
    else if(cfg_.fp32)
        hipLaunchKernelGGL(k_mas_invert<float>, dim3(nbanks), dim3(64), 0, s, B_.data(),
                           inv32_.data());
    else
        hipLaunchKernelGGL(k_mas_invert<double>, dim3(nbanks), dim3(64), 0, s, B_.data(),
                           inv64_.data());
    if(prof)
        prof->end("mas_invert", s);
    SG_LAUNCH_CHECK();
}

template <typename T, typename V>
static void mas_apply_t(Profiler* prof, int nparts, int L, int nv, const std::vector<int>& lvl,
                        const int* vert_of_slot, const int* slot_of, const int* chain,
                        const int* going_next, const int* bank_first, const int* bank_ncl,
                        const int* child_bank, const T* inv, T* R, T* Z, V* x, V* r, const V* p,
                        const V* Ap, const double* rz_prev, const double* pap, V* z,
                        double* rz_out, double* zero0, double* zero1, const int* done,
                        hipStream_t s, bool defer_rz)
{
    // Plan: separate launches for levels 1 .. lt-1 (each stores the next
    // level's R), then one k_mas_coarse launch for the tail lt .. L-1, whose
    // workgroups load R_lt and restrict the levels above in LDS.  (Summing R_lt
    // from level lt-1 inside the tail launch costs two dependent gathers,
    // more than the extra launch.)
    int lt = L;
    if(L > 1)
    {
        for(int l = 1; l < L; l++)
        {
            int nbk  = (lvl[2 * l] + BANK - 1) / BANK;
            int span = lvl[2 * (L - 1) + 1] + lvl[2 * (L - 1)] - lvl[2 * l + 1];
            if(nbk <= kTailBanks && span <= kTailNodes)
            {
                lt = l;
                break;
            }
        }
    }
    auto ev = prof ? prof->kernel_events("mas_apply_l0")
                   : std::pair<hipEvent_t, hipEvent_t>{nullptr, nullptr};
    hipExtLaunchKernelGGL((k_mas_l0<T, V>), dim3(nparts), dim3(64), 0, s, ev.first, ev.second, 0,
                          vert_of_slot, x, r, p, Ap, rz_prev, pap, inv, going_next, bank_first,
                          bank_ncl, R, Z, L > 1 ? 1 : 0, done);
    SG_LAUNCH_CHECK();
    if(L > 1)
    {
        for(int l = 1; l < lt; l++)
        {
            int count = lvl[2 * l], off = lvl[2 * l + 1];
            int nbk   = (count + BANK - 1) / BANK;
            hipLaunchKernelGGL(k_mas_lx<T>, dim3(nbk), dim3(64), 0, s, off / BANK, off, count, inv,
                               going_next, bank_first, bank_ncl, R, Z, l + 1 < L ? 1 : 0, done);
            SG_LAUNCH_CHECK();
        }
        SG_CHECK(L <= 24, -1, "MAS: more than 24 levels");
        LevelTable lv{};
        for(int q = 0; q < L; q++)
        {
            lv.count[q] = lvl[2 * q];
            lv.off[q]   = lvl[2 * q + 1];
        }
        if(lt < L)
        {
            int span = lvl[2 * (L - 1) + 1] + lvl[2 * (L - 1)] - lvl[2 * lt + 1];
            int nbt  = (span + BANK - 1) / BANK;
            int ntw  = (nbt + kCoarseWaves - 1) / kCoarseWaves;
            hipLaunchKernelGGL(k_mas_coarse<T>, dim3(ntw), dim3(kTailThreads), 0, s, lv, L, lt,
                               inv, going_next, child_bank, R, Z, done);
            SG_LAUNCH_CHECK();
        }
    }
    SG_CHECK(grid_for(nv, 256) <= (unsigned)kSetMaxBlocks, SGIPC_ECAPACITY,
             "MAS: system too large for the fixed-order reductions");
    auto collect = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid_for(nv, 256)), dim3(256), 0, s, nv, L, slot_of, chain,
                           Z, r, z, rz_out, zero0, zero1, done);
    };
    static_assert(kChainFast == 8, "collect dispatch below covers NL = 0 .. 8");
    switch(L - 1)
    {
    case 0: collect(k_mas_collect<T, V, 0>); break;
    case 1: collect(k_mas_collect<T, V, 1>); break;
    case 2: collect(k_mas_collect<T, V, 2>); break;
    case 3: collect(k_mas_collect<T, V, 3>); break;
    case 4: collect(k_mas_collect<T, V, 4>); break;
    case 5: collect(k_mas_collect<T, V, 5>); break;
    case 6: collect(k_mas_collect<T, V, 6>); break;
    case 7: collect(k_mas_collect<T, V, 7>); break;
    case 8: collect(k_mas_collect<T, V, 8>); break;
    default: collect(k_mas_collect<T, V, -1>); break;
    }
    SG_LAUNCH_CHECK();
    if(rz_out && !defer_rz)
        set_finalize(rz_out, (int)grid_for(nv, 256), done, s);
}

template <typename V>
void Mas::apply_fused_t(V* x, V* r, const V* p, const V* Ap, const double* rz_prev,
                        const double* pap, V* z, double* rz_out, double* zero0, double* zero1,
                        const int* done, hipStream_t s, bool defer_rz)
{
    SG_CHECK(built_, -5, "MAS apply before build");
    if(cfg_.fp32)
        mas_apply_t<float, V>(prof, nparts_, L_, nv_, h_lvl_, vert_of_slot_.data(),
                              slot_of_.data(), chain_.data(), going_next_.data(),
                              bank_first_.data(), bank_ncl_.data(), child_bank_.data(),
                              inv32_.data(), R32_.data(),
                              Z32_.data(), x, r, p, Ap, rz_prev, pap, z, rz_out, zero0, zero1,
                              done, s, defer_rz);
    else
        mas_apply_t<double, V>(prof, nparts_, L_, nv_, h_lvl_, vert_of_slot_.data(),
                               slot_of_.data(), chain_.data(), going_next_.data(),
                               bank_first_.data(), bank_ncl_.data(), child_bank_.data(),
                               inv64_.data(), R64_.data(),
                               Z64_.data(), x, r, p, Ap, rz_prev, pap, z, rz_out, zero0, zero1,
                               done, s, defer_rz);
}

void Mas::apply_fused(double* x, double* r, const double* p, const double* Ap,
                      const double* rz_prev, const double* pap, double* z, double* rz_out,
                      double* zero0, double* zero1, const int* done, hipStream_t s,
                      bool defer_rz)
{
    apply_fused_t<double>(x, r, p, Ap, rz_prev, pap, z, rz_out, zero0, zero1, done, s, defer_rz);
}

void Mas::apply_fused(float* x, float* r, const float* p, const float* Ap,
                      const double* rz_prev, const double* pap, float* z, double* rz_out,
                      double* zero0, double* zero1, const int* done, hipStream_t s,
                      bool defer_rz)
{
    apply_fused_t<float>(x, r, p, Ap, rz_prev, pap, z, rz_out, zero0, zero1, done, s, defer_rz);
}

void Mas::apply(const double* r, double* z, hipStream_t s)
{
    apply_fused(nullptr, const_cast<double*>(r), nullptr, nullptr, nullptr, nullptr, z, nullptr,
                nullptr, nullptr, nullptr, s);
}

std::vector<int> Mas::chain_host(hipStream_t s) const
{
    std::vector<int> h((size_t)nv_ * (L_ - 1));
    chain_.download(h.data(), h.size(), s);
    SG_HIP(hipStreamSynchronize(s));
    return h;
}

}  // namespace sgipc
