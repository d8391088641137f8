// linsys.hip -- BSR SpMV, residual, block-Jacobi and the PCG driver.
//
// PCG follows the reference's iteration exactly (pcg_solver.cu:171-304):
//   x = 0, r = b, z = M r, p = z;  for k = 1..: Ap = A p, a = rz/(p.Ap),
//   x += a p, r -= a Ap, (every 10th: r = b - A x), z = M r, rz' = r.z,
//   (every 10th: stop if rz' <= rel^2 rz0 or <= abs^2), p = z + (rz'/rz) p.
// Unlike the reference (two device->host copies per iteration, :145) every
// scalar stays on the device: pAp[k] and rz[k] are summed in a fixed order
// into slot sets (common.h) and the consuming kernels derive a and b from
// them; the host reads back only at the convergence test.
#include "linsys.h"

#include "../../include/sgipc.h"

#include <cstdlib>
#include <cstring>

namespace sgipc
{

void throw_hip(hipError_t e, const char* what, const char* file, int line)
{
    char buf[512];
    snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e),
             what, file, line);
    throw Error(-2, buf);
}

void Bsr::set_pattern(int nb_, const std::vector<int>& rp, const std::vector<int>& cl,
                      hipStream_t s)
{
    nb       = nb_;
    nnzb     = rp[nb_];
    h_rowptr = rp;
    h_cols   = cl;
    std::vector<int> rw(nnzb), dg(nb, -1);
    for(int i = 0; i < nb; i++)
        for(int k = rp[i]; k < rp[i + 1]; k++)
        {
            rw[k] = i;
            if(cl[k] == i)
                dg[i] = k;
        }
    for(int i = 0; i < nb; i++)
        SG_CHECK(dg[i] >= 0, -1, "BSR pattern: a block row has no diagonal block");
    rowptr.upload(rp, s);
    cols.upload(cl, s);
    rows.upload(rw, s);
    diag.upload(dg, s);
    vals.resize((size_t)9 * nnzb);
    SG_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------- contact CSR
// Counting sort of the live COO entries by block row: histogram, one-block
// exclusive scan, scatter.  Entries of one row keep no particular order.
__global__ void k_coo_hist(const unsigned* __restrict__ count, unsigned cap,
                           const int* __restrict__ r, int* __restrict__ hist)
{
    unsigned n = *count;
    n          = n < cap ? n : cap;
    unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t < n)
        atomicAdd(hist + r[t], 1);
}

// Exclusive scan of a[0..n) into out[0..n], out[n] = total, in three
// launches: per-tile sums (1024 elements per workgroup), one workgroup scans
// the tile sums, every tile scans itself and adds its offset.
constexpr int kScanTile = 1024;

__device__ __forceinline__ int block_incl_scan(int v, int* ws)  // 1024 threads
{
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    for(int o = 1; o < 64; o <<= 1)
    {
        int u = __shfl_up(v, o, 64);
        if(lane >= o)
            v += u;
    }
    if(lane == 63)
        ws[w] = v;
    __syncthreads();
    if(threadIdx.x < 16)
    {
        int x = ws[threadIdx.x];
        for(int o = 1; o < 16; o <<= 1)
        {
            int u = __shfl_up(x, o, 16);
            if((int)threadIdx.x >= o)
                x += u;
        }
        ws[threadIdx.x] = x;
    }
    __syncthreads();
    return v + (w > 0 ? ws[w - 1] : 0);
}

__global__ void __launch_bounds__(kScanTile) k_scan_tiles(int n, const int* __restrict__ a,
                                                          int* __restrict__ tsum)
{
    __shared__ int ws[16];
    int i = blockIdx.x * kScanTile + threadIdx.x;
    int v = block_incl_scan(i < n ? a[i] : 0, ws);
    if(threadIdx.x == kScanTile - 1)
        tsum[blockIdx.x] = v;
}

__global__ void __launch_bounds__(kScanTile) k_scan_tile_sums(int nt, int* __restrict__ tsum)
{
    __shared__ int ws[16];
    __shared__ int carry;
    if(threadIdx.x == 0)
        carry = 0;
    __syncthreads();
    for(int base = 0; base < nt; base += kScanTile)
    {
        int i = base + threadIdx.x;
        int v = i < nt ? tsum[i] : 0;
        int s = block_incl_scan(v, ws);
        int c = carry;
        if(i < nt)
            tsum[i] = c + s - v;  // exclusive offset of tile i
        __syncthreads();
        if(threadIdx.x == kScanTile - 1)
            carry = c + s;
        __syncthreads();
    }
    if(threadIdx.x == 0)
        tsum[nt] = carry;
}

__global__ void __launch_bounds__(kScanTile) k_scan_apply(int n, const int* __restrict__ a,
                                                          const int* __restrict__ tsum,
                                                          int* __restrict__ out, int nt)
{
    __shared__ int ws[16];
    int i = blockIdx.x * kScanTile + threadIdx.x;
    int v = i < n ? a[i] : 0;
    int s = block_incl_scan(v, ws);
    if(i < n)
        out[i] = tsum[blockIdx.x] + s - v;
    if(blockIdx.x == 0 && threadIdx.x == 0)
        out[n] = tsum[nt];
}

__global__ void k_coo_scatter(const unsigned* __restrict__ count, unsigned cap,
                              const int* __restrict__ r, const int* __restrict__ c,
                              const double* __restrict__ v, const int* __restrict__ rp,
                              int* __restrict__ fill, int* __restrict__ rc,
                              double* __restrict__ rv)
{
    unsigned n = *count;
    n          = n < cap ? n : cap;
    unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= n)
        return;
    int row = r[t];
    int pos = rp[row] + atomicAdd(fill + row, 1);
    rc[pos] = c[t];
    for(int q = 0; q < 9; q++)
        rv[9 * (size_t)pos + q] = v[9 * (size_t)t + q];
}

// ------------------------------------------------------------- merge
// Contact blocks mapped to affine-body nodes repeat the same (row, col) many
// times (every contact between two bodies lands on the same 16 node pairs).
// The contact kernels therefore insert their blocks into an open-addressing
// table keyed by (row, col) (BlockSink::table_add): equal keys are summed on
// insertion with FP64 atomics.  finish() turns the claimed slots into the
// entry arrays and clears them for the next assembly.  This is the merge of
// the reference's triplet converter (converter.cu:23-189, radix sort +
// segmented reduction) done without sorting.

// claimed slots -> entries (r, c, v); slots cleared
__global__ void k_coo_unhash(const unsigned* __restrict__ nlist, const int* __restrict__ list,
                             unsigned long long* __restrict__ keys, double* __restrict__ vals,
                             int* __restrict__ r, int* __restrict__ c, double* __restrict__ v,
                             unsigned* __restrict__ count)
{
    unsigned n = *nlist;
    unsigned u = blockIdx.x * blockDim.x + threadIdx.x;
    if(u == 0)
        *count = n;
    if(u >= n)
        return;
    int                h   = list[u];
    unsigned long long key = keys[h];
    r[u]                   = (int)(key >> 32);
    c[u]                   = (int)(key & 0xffffffffu);
#pragma unroll
    for(int q = 0; q < 9; q++)
    {
        v[9 * (size_t)u + q]      = vals[9 * (size_t)h + q];
        vals[9 * (size_t)h + q] = 0.0;
    }
    keys[h] = kEmptyKey;
}

void Coo::reserve(int nrows, hipStream_t s)
{
    nb = nrows;
    count.resize(1);
    rp.resize(nrows + 1);
    fill.resize(nrows + 1);
    tiles.resize((nrows + 1023) / 1024 + 2);
    hcount.resize(1);
    hfail.resize(1);
    // initial table size: 64 K slots, or SGIPC_CONTACT_TABLE_SLOTS (a power
    // of two >= 16; small values exercise the grow-and-reassemble path)
    size_t      slots = size_t(1) << 16;
    const char* env   = getenv("SGIPC_CONTACT_TABLE_SLOTS");
    if(env)
    {
        long long v = atoll(env);
        SG_CHECK(v >= 16 && (v & (v - 1)) == 0, -1,
                 "SGIPC_CONTACT_TABLE_SLOTS must be a power of two >= 16");
        slots = (size_t)v;
    }
    hcap = 0;
    grow_table(slots, s);
    ensure_entries(1 << 15);
    count.zero(s);
    hcount.zero(s);
    hfail.zero(s);
}

void Coo::grow_table(size_t slots, hipStream_t s)
{
    if(slots <= hcap)
        return;
    hkeys.resize(slots);
    hvals.resize(9 * slots);
    hlist.resize(slots);
    SG_HIP(hipMemsetAsync(hkeys.data(), 0xff, slots * sizeof(unsigned long long), s));
    SG_HIP(hipMemsetAsync(hvals.data(), 0, 9 * slots * sizeof(double), s));
    hcap = slots;
}

void Coo::ensure_entries(int n)
{
    if(n <= cap)
        return;
    int c2 = cap > 0 ? cap : 1;
    while(c2 < n)
        c2 *= 2;
    r.resize(c2);
    c.resize(c2);
    v.resize(9 * (size_t)c2);
    rc.resize(c2);
    rv.resize(9 * (size_t)c2);
    cap = c2;
}

void Coo::begin(hipStream_t s)
{
    SG_HIP(hipMemsetAsync(hcount.data(), 0, sizeof(unsigned), s));
    SG_HIP(hipMemsetAsync(hfail.data(), 0, sizeof(int), s));
    SG_HIP(hipMemsetAsync(count.data(), 0, sizeof(unsigned), s));
    rows_valid = false;
}

bool Coo::finish(hipStream_t s)
{
    unsigned n = 0;
    int      fail = 0;
    SG_HIP(hipMemcpyAsync(&n, hcount.data(), sizeof(unsigned), hipMemcpyDeviceToHost, s));
    SG_HIP(hipMemcpyAsync(&fail, hfail.data(), sizeof(int), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    if(fail)
    {
        // the keys of this attempt stay in the old table: allocate afresh
        size_t want = 4 * hcap;
        hcap        = 0;
        grow_table(want, s);
        return false;
    }
    ensure_entries((int)n);
    if(n > 0)
        hipLaunchKernelGGL(k_coo_unhash, dim3(grid_for(n, 256)), dim3(256), 0, s, hcount.data(),
                           hlist.data(), hkeys.data(), hvals.data(), r.data(), c.data(),
                           v.data(), count.data());
    SG_LAUNCH_CHECK();
    live = n;
    // keep the table at most half full for the next assembly (the slots
    // of this one were cleared above, so a fresh larger table is equivalent)
    if(2 * (size_t)n > hcap)
    {
        size_t want = hcap;
        while(want < 4 * (size_t)n)
            want <<= 1;
        hcap = 0;
        grow_table(want, s);
    }
    return true;
}

void Coo::build_rows(hipStream_t s)
{
    SG_CHECK(nb > 0, -5, "contact CSR: reserve() not called");
    SG_HIP(hipMemsetAsync(fill.data(), 0, (nb + 1) * sizeof(int), s));
    const long long ent = live ? live : (unsigned)cap;  // entries to visit
    hipLaunchKernelGGL(k_coo_hist, dim3(grid_for(ent, 256)), dim3(256), 0, s, count.data(),
                       (unsigned)cap, r.data(), fill.data());
    SG_LAUNCH_CHECK();
    int nt = (nb + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kScanTile), 0, s, nb, fill.data(),
                       tiles.data());
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(kScanTile), 0, s, nt, tiles.data());
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(kScanTile), 0, s, nb, fill.data(),
                       tiles.data(), rp.data(), nt);
    SG_LAUNCH_CHECK();
    SG_HIP(hipMemsetAsync(fill.data(), 0, (nb + 1) * sizeof(int), s));
    hipLaunchKernelGGL(k_coo_scatter, dim3(grid_for(ent, 256)), dim3(256), 0, s, count.data(),
                       (unsigned)cap, r.data(), c.data(), v.data(), rp.data(), fill.data(),
                       rc.data(), rv.data());
    SG_LAUNCH_CHECK();
    rows_valid = true;
}

// ------------------------------------------------------------------ SpMV
// 4 lanes per block row (measured against 8 and 2: 12.95 / 13.25 / 15.3 us
// on a 46,656-row system with 12.5 blocks per row, 48.4 / 51.2 / 54.7 us on
// 175,273 rows with 14.5); a row's blocks are contiguous, so the 4 lanes of a
// group stream 288 contiguous bytes per step (144 in FP32).  The static
// pattern and the row-sorted contact blocks of the row are traversed by the
// same lanes, so y is written once per row without atomics.  V is the
// precision of the matrix, the vectors and the row sums.
//
// MODE_PLAIN: y = A x            (+ optional x.y into a slot set)
// MODE_RESID: y = b - A x
// MODE_PCG:   the PCG direction is formed on the fly: for every column j read,
//             p_j = z_j + beta p_prev_j (beta = rz[k-1] / rz[k-2], slot sets),
//             or p_j = z_j on the first iteration; the row owner stores p_i,
//             y = A p and p.y is accumulated into a slot set.  Every read of
//             p_j evaluates the same fma, so all rows see identical values.
constexpr int kRowLanes = 4;
enum
{
    MODE_PLAIN = 0,
    MODE_RESID = 1,
    MODE_PCG   = 2
};

// PCG control (MODE_PCG, MODE_RESID): the kernel returns at once when
// ctl[0] != 0 (the solve has stopped).  With q_set != null the previous
// iteration was a convergence check: every workgroup evaluates the stop test
// q <= thr (q: the slot-set sum of r.z or r.r) and, if it holds or r.z is
// not finite, returns; workgroup 0 records the stopping iteration check_k in
// ctl[0] (ctl[1] = 1 on breakdown).  The r.z sum is kept in ctl_d[0].
struct PcgCtl
{
    int*          ctl      = nullptr;
    double*       ctl_d    = nullptr;
    const double* q_set    = nullptr;
    const double* rz_check = nullptr;
    double        thr      = 0.0;
    double        div_thr  = 0.0;  // > 0: q above it stops the solve as diverged (ctl[1] = 2)
    int           check_k  = 0;
};

template <int MODE, typename V>
__global__ void __launch_bounds__(256)
    k_spmv(int nb, const int* __restrict__ rowptr, const int* __restrict__ cols,
           const V* __restrict__ vals, const int* __restrict__ rp2,
           const int* __restrict__ cols2, const V* __restrict__ vals2,
           const V* __restrict__ x, const V* __restrict__ pprev, V* __restrict__ pcur,
           const V* __restrict__ b, V* __restrict__ y, double* rz_a, const double* rz_b,
           int rz_parts, double* dot, PcgCtl pc)
{
    if(pc.ctl && pc.ctl[0])
        return;
    // contiguous row ranges per XCD: a row's neighbours (x gathers) are
    // mostly rows of the same range, so they hit that XCD's L2
    int  gid  = (int)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    int  row  = gid / kRowLanes;
    int  g    = gid % kRowLanes;
    bool live = row < nb;
    // independent loads first: row extents (static and contact rows)
    int k0 = 0, k1 = 0, q0 = 0, q1 = 0;
    if(live)
    {
        k0 = rowptr[row];
        k1 = rowptr[row + 1];
        if(rp2)
        {
            q0 = rp2[row];
            q1 = rp2[row + 1];
        }
    }
    const V* pp = (MODE == MODE_PCG && pprev) ? pprev : x;
    // the row's own p_i (stored by the row owner, and its p.Ap term): raw
    // loads issued before the passes, combined after them
    const bool own = live && g == 0 && (MODE == MODE_PCG || (MODE == MODE_PLAIN && dot));
    V          ox[3] = {0, 0, 0}, op[3] = {0, 0, 0};
    if(own)
    {
#pragma unroll
        for(int c = 0; c < 3; c++)
        {
            ox[c] = x[3 * (size_t)row + c];
            if(MODE == MODE_PCG)
                op[c] = pp[3 * (size_t)row + c];
        }
    }
    // FP64: the first pass of the static row (blocks k0 + g and k0 + g +
    // kRowLanes), column ids and blocks requested before the r.z total below
    // is waited for.  (FP32: the extra registers cost more occupancy than the
    // overlap gains; measured 26.4 against 23.7 us on a 175k-row system.)
    constexpr bool kHoist = sizeof(V) == 8;
    const int      kf     = k0 + g;
    const bool     has1   = kHoist && live && kf < k1;
    const bool two1 = has1 && kf + kRowLanes < k1;
    int        ca1 = 0, cb1 = 0;
    V          A1[9], B1[9];
    if(has1)
    {
        const int kk = two1 ? kf + kRowLanes : kf;
        ca1          = cols[kf];
        cb1          = cols[kk];
        const V* pa  = vals + 9 * (size_t)kf;
        const V* pb  = vals + 9 * (size_t)kk;
#pragma unroll
        for(int q = 0; q < 9; q++)
        {
            A1[q] = pa[q];
            B1[q] = pb[q];
        }
    }
    // rz_parts > 0: the set rz_a holds the preconditioner's rz_parts
    // workgroup partials, not yet finalized.  One wave per workgroup sums them
    // in a fixed order (every workgroup gets the same bits) and workgroup 0
    // stores the total in slot 0 for the launches that follow.
    double rz_new = 0.0;
    if(MODE == MODE_PCG && rz_parts > 0)
    {
        __shared__ double rz_sh;
        if(threadIdx.x < 64)
        {
            double t = slot_set_partials_sum(rz_a, rz_parts);
            if(threadIdx.x == 0)
                rz_sh = t;
        }
        __syncthreads();
        rz_new = rz_sh;
        if(blockIdx.x == 0 && threadIdx.x == 0)
            rz_a[0] = rz_new;
    }
    if(MODE == MODE_PCG && pc.q_set)
    {
        double q   = (rz_parts > 0 && pc.q_set == rz_a) ? rz_new : slot_set_sum(pc.q_set);
        double rzc = (rz_parts > 0 && pc.rz_check == rz_a) ? rz_new : slot_set_sum(pc.rz_check);
        bool   bad = !(rzc == rzc) || !(q == q);
        bool   div = pc.div_thr > 0.0 && q > pc.div_thr;
        if(blockIdx.x == 0 && threadIdx.x == 0)
        {
            pc.ctl_d[0] = rzc;
            pc.ctl_d[1] = q;  // the tested quantity, read back by the host
        }
        if(q <= pc.thr || bad || div)
        {
            if(blockIdx.x == 0 && threadIdx.x == 0)
            {
                pc.ctl[1] = bad ? 1 : (div ? 2 : 0);
                pc.ctl[0] = pc.check_k;
            }
            return;
        }
    }
    // MODE_PCG: p_j = z_j + beta p_prev_j; on the first iteration (no p_prev)
    // beta = 0 and p_prev aliases z, so fma(0, z_j, z_j) = z_j and every
    // gather takes the same branch-free path
    V beta = 0;
    if(MODE == MODE_PCG && pprev)
    {
        double sa = rz_parts > 0 ? 0.0 : rz_a[__lane_id() * kSlotStride];
        double sb = rz_b[__lane_id() * kSlotStride];
        sa        = rz_parts > 0 ? rz_new : wave_sum(sa);
        beta      = (V)(sa / wave_sum(sb));  // slot_set_sum of both sets
    }
    auto fetch = [&](int j, V& x0, V& x1, V& x2) {
        const V* xj = x + 3 * (size_t)j;
        if(MODE == MODE_PCG)
        {
            const V* pj = pp + 3 * (size_t)j;
            x0          = fma(beta, pj[0], xj[0]);
            x1          = fma(beta, pj[1], xj[1]);
            x2          = fma(beta, pj[2], xj[2]);
        }
        else
        {
            x0 = xj[0];
            x1 = xj[1];
            x2 = xj[2];
        }
    };
    V a0 = 0, a1 = 0, a2 = 0;
    // A lane takes blocks k, k + 8, k + 16, ... of its row, two per pass: both
    // column ids and both blocks are loaded together, then both gathers, so
    // a pass costs two memory round trips.  Accumulation order is the
    // block order k, k + 8, ... as with one block per pass.
    auto row_pass = [&](const int* __restrict__ cl, const V* __restrict__ vl, int lo, int hi) {
        for(int k = lo + g; k < hi; k += 2 * kRowLanes)
        {
            const int  kb  = k + kRowLanes;
            const bool two = kb < hi;
            const int  kk  = two ? kb : k;  // in range either way
            const int  ca  = cl[k];
            const int  cb  = cl[kk];
            V          A[9], B[9];
            const V*   pa = vl + 9 * (size_t)k;
            const V*   pb = vl + 9 * (size_t)kk;
#pragma unroll
            for(int q = 0; q < 9; q++)
            {
                A[q] = pa[q];
                B[q] = pb[q];
            }
            // keep the block loads ahead of the gathers (the scheduler would
            // otherwise issue them after the column ids have returned)
            __builtin_amdgcn_sched_barrier(0);
            V x0, x1, x2, y0, y1, y2;
            fetch(ca, x0, x1, x2);
            fetch(cb, y0, y1, y2);
#pragma unroll
            for(int q = 0; q < 9; q++)
                B[q] = two ? B[q] : (V)0;  // a zero block adds exactly +0
            a0 += A[0] * x0 + A[1] * x1 + A[2] * x2;
            a1 += A[3] * x0 + A[4] * x1 + A[5] * x2;
            a2 += A[6] * x0 + A[7] * x1 + A[8] * x2;
            a0 += B[0] * y0 + B[1] * y1 + B[2] * y2;
            a1 += B[3] * y0 + B[4] * y1 + B[5] * y2;
            a2 += B[6] * y0 + B[7] * y1 + B[8] * y2;
        }
    };
    // the first static pass, with the blocks loaded above (the order of
    // the products is the one row_pass would give)
    if(has1)
    {
        V x0, x1, x2, y0, y1, y2;
        fetch(ca1, x0, x1, x2);
        fetch(cb1, y0, y1, y2);
#pragma unroll
        for(int q = 0; q < 9; q++)
            B1[q] = two1 ? B1[q] : (V)0;  // a zero block adds exactly +0
        a0 += A1[0] * x0 + A1[1] * x1 + A1[2] * x2;
        a1 += A1[3] * x0 + A1[4] * x1 + A1[5] * x2;
        a2 += A1[6] * x0 + A1[7] * x1 + A1[8] * x2;
        a0 += B1[0] * y0 + B1[1] * y1 + B1[2] * y2;
        a1 += B1[3] * y0 + B1[4] * y1 + B1[5] * y2;
        a2 += B1[6] * y0 + B1[7] * y1 + B1[8] * y2;
    }
    if(live)
    {
        row_pass(cols, vals, kHoist ? k0 + 2 * kRowLanes : k0, k1);
        if(rp2)
            row_pass(cols2, vals2, q0, q1);
    }
#pragma unroll
    for(int o = kRowLanes / 2; o > 0; o >>= 1)
    {
        a0 += __shfl_xor(a0, o, kRowLanes);
        a1 += __shfl_xor(a1, o, kRowLanes);
        a2 += __shfl_xor(a2, o, kRowLanes);
    }
    double d = 0.0;
    if(live && g == 0)
    {
        size_t i = 3 * (size_t)row;
        if(MODE == MODE_RESID)
        {
            y[i + 0] = b[i + 0] - a0;
            y[i + 1] = b[i + 1] - a1;
            y[i + 2] = b[i + 2] - a2;
        }
        else
        {
            y[i + 0] = a0;
            y[i + 1] = a1;
            y[i + 2] = a2;
            if(own)
            {
                V x0 = ox[0], x1 = ox[1], x2 = ox[2];
                if(MODE == MODE_PCG)
                {
                    x0 = fma(beta, op[0], ox[0]);
                    x1 = fma(beta, op[1], ox[1]);
                    x2 = fma(beta, op[2], ox[2]);
                }
                if(MODE == MODE_PCG)
                {
                    pcur[i + 0] = x0;
                    pcur[i + 1] = x1;
                    pcur[i + 2] = x2;
                }
                d = (double)(x0 * a0 + x1 * a1 + x2 * a2);
            }
        }
    }
    if(MODE == MODE_PCG || (MODE == MODE_PLAIN && dot))
        slot_set_add(dot, d);
}

// matrix values of the requested precision (static pattern, contact rows)
template <typename V>
struct MatVals;
template <>
struct MatVals<double>
{
    static const double* a(const SysView& S) { return S.A->vals.data(); }
    static const double* c(const Coo* c) { return c->rv.data(); }
};
template <>
struct MatVals<float>
{
    static const float* a(const SysView& S) { return S.A->vals32.data(); }
    static const float* c(const Coo* c) { return c->rv32.data(); }
};

template <typename V>
static inline void contact_rows(const SysView& S, const int*& rp2, const int*& c2, const V*& v2)
{
    rp2 = nullptr;
    c2  = nullptr;
    v2  = nullptr;
    if(S.coo && S.coo->cap > 0)
    {
        SG_CHECK(S.coo->rows_valid, -5, "contact CSR is stale: build_rows() after assembly");
        rp2 = S.coo->rp.data();
        c2  = S.coo->rc.data();
        v2  = MatVals<V>::c(S.coo);
    }
}

template <int MODE, typename V>
static void launch_spmv(const SysView& S, const V* x, const V* pprev, V* pcur, const V* b, V* y,
                        double* rz_a, const double* rz_b, double* dot, const PcgCtl& pc,
                        hipStream_t s, std::pair<hipEvent_t, hipEvent_t> ev = {nullptr, nullptr},
                        int rz_parts = 0)
{
    const Bsr& A = *S.A;
    const int* rp2;
    const int* c2;
    const V*   v2;
    contact_rows<V>(S, rp2, c2, v2);
    SG_CHECK(grid_for((long long)A.nb * kRowLanes, 256) <= (unsigned)kSetMaxBlocks, SGIPC_ECAPACITY,
             "system too large for the fixed-order reductions (block rows > 4M)");
    hipExtLaunchKernelGGL((k_spmv<MODE, V>), dim3(grid_for((long long)A.nb * kRowLanes, 256)),
                          dim3(256), 0, s, ev.first, ev.second, 0, A.nb, A.rowptr.data(),
                          A.cols.data(), MatVals<V>::a(S), rp2, c2, v2, x, pprev, pcur, b, y, rz_a,
                          rz_b, rz_parts, dot, pc);
    SG_LAUNCH_CHECK();
    if(dot)
        set_finalize(dot, (int)grid_for((long long)A.nb * kRowLanes, 256), pc.ctl, s);
}

void spmv(const SysView& S, const double* x, double* y, double* dot, hipStream_t s)
{
    launch_spmv<MODE_PLAIN, double>(S, x, nullptr, nullptr, nullptr, y, nullptr, nullptr, dot,
                                    PcgCtl{}, s);
}

void residual(const SysView& S, const double* b, const double* x, double* r, hipStream_t s)
{
    launch_spmv<MODE_RESID, double>(S, x, nullptr, nullptr, b, r, nullptr, nullptr, nullptr,
                                    PcgCtl{}, s);
}

template <typename V>
__global__ void __launch_bounds__(256) k_dot_slots(int n, const V* __restrict__ a,
                                                   const V* __restrict__ b, double* out,
                                                   const int* __restrict__ done)
{
    if(done && *done)
        return;
    int    i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.0;
    for(; i < n; i += gridDim.x * blockDim.x)
        d += (double)a[i] * (double)b[i];
    slot_set_add(out, d);
}

template <typename V>
static void dot_slots_ctl(int n, const V* a, const V* b, double* out, const int* done,
                          hipStream_t s)
{
    unsigned g = grid_for(n, 256);
    g          = g > 1024 ? 1024 : g;
    hipLaunchKernelGGL(k_dot_slots<V>, dim3(g), dim3(256), 0, s, n, a, b, out, done);
    SG_LAUNCH_CHECK();
    set_finalize(out, (int)g, done, s);
}

void dot_slots(int n, const double* a, const double* b, double* out, hipStream_t s)
{
    dot_slots_ctl<double>(n, a, b, out, nullptr, s);
}

__global__ void __launch_bounds__(kFinalizeThreads) k_set_finalize(double* set, int nblocks,
                                                                   const int* __restrict__ done)
{
    if(done && *done)
        return;
    slot_set_finalize(set, nblocks);
}

void set_finalize(double* set, int nblocks, const int* done, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_finalize, dim3(1), dim3(kFinalizeThreads), 0, s, set, nblocks, done);
    SG_LAUNCH_CHECK();
}

// clear a slot set in stream order (a kernel, so the sequence stays free of
// memset nodes if it is ever captured into a graph)
__global__ void k_clear_set(double* set, const int* __restrict__ done)
{
    if(done && *done)
        return;
    slot_set_clear(set, threadIdx.x);
}

double slot_set_sum_host(const double* h)
{
    // the device sums lane values with a butterfly: after the step with
    // offset o every lane holds v[l] + v[l ^ o]; lane 0's result is
    // reproduced here step by step for all lanes
    double v[kSlots];
    for(int l = 0; l < kSlots; l++)
        v[l] = h[l * kSlotStride];
    for(int o = 32; o > 0; o >>= 1)
    {
        double u[kSlots];
        for(int l = 0; l < kSlots; l++)
            u[l] = v[l] + v[l ^ o];
        for(int l = 0; l < kSlots; l++)
            v[l] = u[l];
    }
    return v[0];
}

double read_slot_set(const double* d, hipStream_t s)
{
    double h[kSlotSet];
    SG_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    return slot_set_sum_host(h);
}

// ------------------------------------------------------------ conversions
template <typename A, typename B>
__global__ void k_convert(long long n, const A* __restrict__ a, B* __restrict__ b)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if(i < n)
        b[i] = (B)a[i];
}

template <typename A, typename B>
static void convert(long long n, const A* a, B* b, hipStream_t s)
{
    if(n <= 0)
        return;
    hipLaunchKernelGGL((k_convert<A, B>), dim3(grid_for(n, 256)), dim3(256), 0, s, n, a, b);
    SG_LAUNCH_CHECK();
}

void refresh_fp32(const SysView& S, hipStream_t s)
{
    Bsr& A = *const_cast<Bsr*>(S.A);
    A.vals32.resize((size_t)9 * A.nnzb);
    convert<double, float>(9ll * A.nnzb, A.vals.data(), A.vals32.data(), s);
    if(S.coo && S.coo->cap > 0)
    {
        Coo& c = *const_cast<Coo*>(S.coo);
        c.rv32.resize(9 * (size_t)c.cap);
        // the live entries are a prefix of the row-sorted copy: convert all
        convert<double, float>(9ll * c.cap, c.rv.data(), c.rv32.data(), s);
    }
}

// ------------------------------------------------------------ block Jacobi
__global__ void k_diag_inverse(int nb, const int* __restrict__ diag,
                               const double* __restrict__ vals, double* __restrict__ dinv)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= nb)
        return;
    const double* a = vals + 9 * (size_t)diag[i];
    double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8],
           c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    double* o  = dinv + 9 * (size_t)i;
    if(det == 0.0)
    {
        for(int k = 0; k < 9; k++)
            o[k] = (k % 4 == 0) ? 1.0 : 0.0;
        return;
    }
    double id = 1.0 / det;
    o[0]      = c00 * id;
    o[1]      = (a[2] * a[7] - a[1] * a[8]) * id;
    o[2]      = (a[1] * a[5] - a[2] * a[4]) * id;
    o[3]      = c01 * id;
    o[4]      = (a[0] * a[8] - a[2] * a[6]) * id;
    o[5]      = (a[2] * a[3] - a[0] * a[5]) * id;
    o[6]      = c02 * id;
    o[7]      = (a[1] * a[6] - a[0] * a[7]) * id;
    o[8]      = (a[0] * a[4] - a[1] * a[3]) * id;
}

// optional update x += a p, r -= a Ap (a = rz[k-1] / pAp[k] from slot sets),
// then z = Dinv r and r.z into a slot set; block 0 clears two later sets
template <typename V>
__global__ void __launch_bounds__(256)
    k_update_diag(int nb, V* x, V* r, const V* p, const V* Ap, const double* rz_prev,
                  const double* pap, const V* __restrict__ dinv, V* z, double* rz_out,
                  double* zero0, double* zero1, const int* __restrict__ done)
{
    if(done && *done)
        return;
    double a = 0.0;
    if(x)
        a = slot_set_sum(rz_prev) / slot_set_sum(pap);
    V      av = (V)a;
    int    i  = blockIdx.x * blockDim.x + threadIdx.x;
    double d  = 0.0;
    if(i < nb)
    {
        V ri[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
        if(x)
        {
            for(int c = 0; c < 3; c++)
            {
                x[3 * i + c] += av * p[3 * i + c];
                ri[c] -= av * Ap[3 * i + c];
                r[3 * i + c] = ri[c];
            }
        }
        const V* D = dinv + 9 * (size_t)i;
        for(int c = 0; c < 3; c++)
        {
            V zc         = D[3 * c] * ri[0] + D[3 * c + 1] * ri[1] + D[3 * c + 2] * ri[2];
            z[3 * i + c] = zc;
            d += (double)(ri[c] * zc);
        }
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

template <typename V>
__global__ void __launch_bounds__(256) k_axpy(int n, V* x, V* r, const V* p, const V* Ap,
                                              const double* rz_prev, const double* pap,
                                              const int* __restrict__ done)
{
    if(done && *done)
        return;
    V   a = (V)(slot_set_sum(rz_prev) / slot_set_sum(pap));
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
    {
        x[i] += a * p[i];
        r[i] -= a * Ap[i];
    }
}

// reliable updates of the mixed mode: x64 += x, x = 0  /  r = (V) r64
template <typename V>
__global__ void k_fold(int n, double* __restrict__ x64, V* __restrict__ x,
                       const int* __restrict__ done)
{
    if(done && *done)
        return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
    {
        x64[i] += (double)x[i];
        x[i] = (V)0;
    }
}

template <typename V>
__global__ void k_round(int n, const double* __restrict__ r64, V* __restrict__ r,
                        const int* __restrict__ done)
{
    if(done && *done)
        return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
        r[i] = (V)r64[i];
}

void Pcg::build_diag(const Bsr& A, hipStream_t s)
{
    dinv_.resize((size_t)9 * A.nb);
    hipLaunchKernelGGL(k_diag_inverse, dim3(grid_for(A.nb, 256)), dim3(256), 0, s, A.nb,
                       A.diag.data(), A.vals.data(), dinv_.data());
    SG_LAUNCH_CHECK();
    if(nbody_ > 0)
    {
        SG_CHECK(body_first_ + 4 * nbody_ == A.nb, -1, "affine-body nodes must end the system");
        binv_.resize((size_t)144 * nbody_);
        launch_body_inverse(nbody_, body_slots_, A.vals.data(), binv_.data(), s);
    }
}

template <>
const double* Pcg::binv<double>() const
{
    return binv_.data();
}
template <>
const float* Pcg::binv<float>() const
{
    return binv32_.data();
}



Pcg::~Pcg()
{
    if(h_ctl_ && !external_)
        (void)hipHostFree(h_ctl_);
    for(auto& e : ev_)
        if(e)
            (void)hipEventDestroy(e);
}

template <>
PcgVecs<double>& Pcg::vecs<double>()
{
    return v64_;
}
template <>
PcgVecs<float>& Pcg::vecs<float>()
{
    return v32_;
}
template <>
const double* Pcg::dinv<double>() const
{
    return dinv_.data();
}
template <>
const float* Pcg::dinv<float>() const
{
    return dinv32_.data();
}

// the global 3x3 block Jacobi on the nodes before the bodies, then the
// bodies' 12x12 blocks (GlobalLinearSystem::apply_preconditioner applies the
// global preconditioner and then the ABD local one over the body range,
// global_linear_system.cu:509-522)
template <typename V>
void Pcg::apply_block_jacobi(int nb, V* x, V* r, const V* p, const V* Ap, const double* rz_prev,
                             const double* pap, V* z, double* rz_out, double* zero0,
                             double* zero1, const int* done, hipStream_t s)
{
    const int n3 = nbody_ > 0 ? body_first_ : nb;
    if(n3 > 0 || zero0 || zero1)
    {
        SG_CHECK(grid_for(n3 > 0 ? n3 : 1, 256) <= (unsigned)kSetMaxBlocks, SGIPC_ECAPACITY,
                 "system too large for the fixed-order reductions");
        hipLaunchKernelGGL(k_update_diag<V>, dim3(grid_for(n3 > 0 ? n3 : 1, 256)), dim3(256), 0,
                           s, n3, x, r, p, Ap, rz_prev, pap, dinv<V>(), z, rz_out, zero0, zero1,
                           done);
        SG_LAUNCH_CHECK();
        if(rz_out)
            set_finalize(rz_out, (int)grid_for(n3 > 0 ? n3 : 1, 256), done, s);
    }
    if(nbody_ > 0)
        launch_update_body<V>(nbody_, body_first_, x, r, p, Ap, rz_prev, pap, binv<V>(), z,
                              rz_out, done, s);
}

void Pcg::apply_diag(int nb, const double* r, double* z, hipStream_t s)
{
    apply_block_jacobi<double>(nb, (double*)nullptr, const_cast<double*>(r), nullptr, nullptr,
                               nullptr, nullptr, z, nullptr, nullptr, nullptr, nullptr, s);
}

size_t Pcg::workspace_doubles(int n)
{
    // reductions, 4 control ints (2 doubles), 4 control doubles, 9 FP64 work
    // vectors and the 3x3 block-Jacobi inverse (9 per block row = 3 n), each
    // region rounded up to 256 bytes
    auto up = [](size_t d) { return (d + 31) / 32 * 32; };
    return up((size_t)(2 * kRing + 4) * kSlotSet) + up(2) + up(4) + 9 * up((size_t)n)
           + up(3 * (size_t)n);
}

void Pcg::bind_workspace(double* base, int n)
{
    auto    up = [](size_t d) { return (d + 31) / 32 * 32; };
    double* p  = base;
    red_.borrow(p, (size_t)(2 * kRing + 4) * kSlotSet);
    p += up((size_t)(2 * kRing + 4) * kSlotSet);
    ctl_.borrow(reinterpret_cast<int*>(p), 4);
    p += up(2);
    ctl_d_.borrow(p, 4);
    p += up(4);
    for(DBuf<double>* b : {&v64_.r, &v64_.z, &v64_.p[0], &v64_.p[1], &v64_.Ap, &v64_.tmp,
                           &v64_.res, &v64_.rhs, &v64_.sol})
    {
        b->borrow(p, (size_t)n);
        p += up((size_t)n);
    }
    v64_.n = n;
    dinv_.borrow(p, 3 * (size_t)n);
    if(h_ctl_ && !external_)
        (void)hipHostFree(h_ctl_);
    h_ctl_    = h_ctl_local_;
    external_ = true;
    n_        = n;
}

void Pcg::ensure(int n)
{
    if(external_)
    {
        SG_CHECK(n == n_, -1, "solver workspace bound for another size");
        return;
    }
    red_.resize((size_t)(2 * kRing + 4) * kSlotSet);
    ctl_.resize(4);
    ctl_d_.resize(4);
    if(!h_ctl_)
    {
        SG_HIP(hipHostMalloc((void**)&h_ctl_, 8 * sizeof(int), hipHostMallocDefault));
        for(auto& e : ev_)
            SG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    (void)n;
}

// Iteration k (1-based) as launched here:
//   spmv_pcg   p_k = z_{k-1} + beta p_{k-1} (beta = rz[k-1]/rz[k-2]; p_1 = z_0),
//              Ap = A p_k, pAp[k]; if k-1 was a check iteration, first the
//              stop test of iteration k-1
//   normal k:  MAS level 0 with x += a p_k, r -= a Ap (a = rz[k-1]/pAp[k]),
//              coarse levels, collect z_k and rz[k]
//   k % every: x, r update; r = b - A x; z_k = M r; rz[k] (and r.r)
// which is the reference's sequence (pcg_solver.cu:171-304) with its
// p = z + beta p step moved into the next iteration's SpMV and its stop test
// evaluated on the device.  Once the test holds every later kernel returns at
// once, so the host never waits inside the loop: it copies the stop flag out
// after each test and reads it one check period later.  Slot sets for
// (pAp, rz) of iteration k live in ring entry k % kRing; the collect of
// iteration k clears entry (k + 2) % kRing, whose last reader was the SpMV
// of iteration k.  V: precision of the matrix, vectors and preconditioner
// application (double, or float for the inner solve of the mixed mode).
template <typename V>
PcgStats Pcg::solve_t(const SysView& S, Precond pc, Mas* mas, const V* b, V* x,
                      const PcgConfig& cfg, hipStream_t s, const double* b64, double* x64)
{
    const int   nb    = S.A->nb;
    const int   n     = 3 * nb;
    int         maxit = cfg.max_iter > 0 ? cfg.max_iter : n;
    int         every = cfg.check_every > 0 ? cfg.check_every : 10;
    PcgVecs<V>& W     = vecs<V>();
    ensure(n);
    W.ensure(n);
    SG_HIP(hipMemsetAsync(red_.data(), 0, red_.size() * sizeof(double), s));
    SG_HIP(hipMemsetAsync(ctl_.data(), 0, ctl_.size() * sizeof(int), s));
    SG_HIP(hipMemsetAsync(ctl_d_.data(), 0, ctl_d_.size() * sizeof(double), s));
    SG_HIP(hipMemsetAsync(x, 0, n * sizeof(V), s));
    SG_HIP(hipMemcpyAsync(W.r.data(), b, n * sizeof(V), hipMemcpyDeviceToDevice, s));
    if(pc == Precond::Diag)
    {
        build_diag(*S.A, s);
        if(sizeof(V) == sizeof(float))
        {
            dinv32_.resize(dinv_.size());
            convert<double, float>((long long)dinv_.size(), dinv_.data(), dinv32_.data(), s);
            if(nbody_ > 0)
            {
                binv32_.resize(binv_.size());
                convert<double, float>((long long)binv_.size(), binv_.data(), binv32_.data(), s);
            }
        }
    }
    const int* done = ctl_.data();

    // z = M r with the optional update; rz[k] -> rz_set(k).  With MAS, the
    // sets rz_set(k), k >= 1, are left as workgroup partials: the SpMV of
    // iteration k + 1 (their first reader) sums them (rz_parts).
    const int rz_parts = pc == Precond::Mas ? mas->collect_blocks() : 0;
    auto precond = [&](bool axpy, int k, const V* p, const int* dn) {
        double* rz_prev = axpy ? rz_set(k - 1) : nullptr;
        double* pap     = axpy ? pap_set(k) : nullptr;
        double* zero0   = pap_set(k + 2);
        double* zero1   = rz_set(k + 2);
        if(pc == Precond::Mas)
            mas->apply_fused(axpy ? x : (V*)nullptr, W.r.data(), p, W.Ap.data(), rz_prev, pap,
                             W.z.data(), rz_set(k), zero0, zero1, dn, s, k >= 1);
        else
            apply_block_jacobi<V>(nb, axpy ? x : (V*)nullptr, W.r.data(), p, W.Ap.data(),
                                  rz_prev, pap, W.z.data(), rz_set(k), zero0, zero1, dn, s);
    };

    const bool reliable = (b64 != nullptr);  // FP64 residual replacement (mixed mode)
    PcgStats   st;
    precond(false, 0, nullptr, nullptr);
    if(!cfg.use_rz)
    {
        if(reliable)
            dot_slots_ctl<double>(n, b64, b64, aux_set(0), nullptr, s);
        else
            dot_slots_ctl<V>(n, W.r.data(), W.r.data(), aux_set(0), nullptr, s);
    }
    double rz0 = read_slot_set(rz_set(0), s);
    double rr0 = cfg.use_rz ? 0.0 : read_slot_set(aux_set(0), s);
    st.rz0     = rz0;
    if(rz0 == 0.0 || !(rz0 == rz0))
    {
        st.iterations = 0;
        st.converged  = (rz0 == 0.0);
        return st;
    }
    double rel2 = cfg.rel_tol * cfg.rel_tol, abs2 = cfg.abs_tol * cfg.abs_tol;
    double q0   = cfg.q_ref > 0.0 ? cfg.q_ref : (cfg.use_rz ? rz0 : rr0);
    double thr  = rel2 * q0 > abs2 ? rel2 * q0 : abs2;

    PcgCtl ctl_only;
    ctl_only.ctl   = ctl_.data();
    ctl_only.ctl_d = ctl_d_.data();
    int  slot = 0, pend_slot = 0;
    bool have_pending = false;
    // mixed mode: the FP32 iteration is abandoned when the tested FP64
    // residual has not halved over cfg.stall_checks consecutive tests (the
    // caller then continues in FP64 from the FP64 iterate)
    double best_q = 0.0;
    int    stall = 0, stalled_at = 0;
    for(int k = 1; k <= maxit; k++)
    {
        PcgCtl     pcx  = ctl_only;
        const bool test = (k - 1 >= 1) && ((k - 1) % every == 0);
        if(test)
        {
            pcx.q_set    = cfg.use_rz ? rz_set(k - 1) : aux_set(1);
            pcx.rz_check = rz_set(k - 1);
            pcx.thr      = thr;
            pcx.div_thr  = reliable ? 10.0 * q0 : 0.0;
            pcx.check_k  = k - 1;
        }
        V*       pcur  = W.p[k & 1].data();
        const V* pprev = (k == 1) ? nullptr : W.p[(k - 1) & 1].data();
        launch_spmv<MODE_PCG, V>(S, W.z.data(), pprev, pcur, nullptr, W.Ap.data(), rz_set(k - 1),
                                 k >= 2 ? rz_set(k - 2) : rz_set(k - 1), pap_set(k), pcx, s,
                                 prof ? prof->kernel_events("spmv")
                                      : std::pair<hipEvent_t, hipEvent_t>{nullptr, nullptr},
                                 k >= 2 ? rz_parts : 0);
        if(test)
        {
            SG_HIP(hipMemcpyAsync(h_ctl_ + 4 * slot, ctl_.data(), 2 * sizeof(int),
                                  hipMemcpyDeviceToHost, s));
            if(reliable && cfg.stall_checks > 0)
                SG_HIP(hipMemcpyAsync(h_ctl_ + 4 * slot + 2, ctl_d_.data() + 1, sizeof(double),
                                      hipMemcpyDeviceToHost, s));
            if(external_)
                SG_HIP(hipStreamSynchronize(s));  // the flag is in pageable memory
            else
                SG_HIP(hipEventRecord(ev_[slot], s));
            if(have_pending)
            {
                if(!external_)
                    SG_HIP(hipEventSynchronize(ev_[pend_slot]));
                if(h_ctl_[4 * pend_slot] != 0)
                    break;
                if(reliable && cfg.stall_checks > 0)
                {
                    double q;
                    memcpy(&q, h_ctl_ + 4 * pend_slot + 2, sizeof(double));
                    if(best_q == 0.0 || q < 0.5 * best_q)
                    {
                        best_q = q;
                        stall  = 0;
                    }
                    else if(++stall >= cfg.stall_checks)
                    {
                        stalled_at = k;  // launches up to k are in the stream
                        break;
                    }
                }
            }
            have_pending = true;
            pend_slot    = slot;
            slot ^= 1;
        }
        if(k == maxit)
            break;  // the last launch only evaluated the final stop test
        if(k % every == 0)
        {
            hipLaunchKernelGGL(k_axpy<V>, dim3(grid_for(n, 256)), dim3(256), 0, s, n, x,
                               W.r.data(), pcur, W.Ap.data(), rz_set(k - 1), pap_set(k), done);
            SG_LAUNCH_CHECK();
            if(reliable)
            {
                // x64 += x; x = 0; r = b - A x64 in FP64, rounded into the
                // working residual (the search direction is kept)
                hipLaunchKernelGGL(k_fold<V>, dim3(grid_for(n, 256)), dim3(256), 0, s, n, x64, x,
                                   done);
                SG_LAUNCH_CHECK();
                launch_spmv<MODE_RESID, double>(S, x64, nullptr, nullptr, b64, v64_.res.data(),
                                                nullptr, nullptr, nullptr, ctl_only, s);
                hipLaunchKernelGGL(k_round<V>, dim3(grid_for(n, 256)), dim3(256), 0, s, n,
                                   v64_.res.data(), W.r.data(), done);
                SG_LAUNCH_CHECK();
            }
            else
                launch_spmv<MODE_RESID, V>(S, x, nullptr, nullptr, b, W.r.data(), nullptr,
                                           nullptr, nullptr, ctl_only, s);
            precond(false, k, pcur, done);
            if(!cfg.use_rz)
            {
                hipLaunchKernelGGL(k_clear_set, dim3(1), dim3(64), 0, s, aux_set(1), done);
                SG_LAUNCH_CHECK();
                if(reliable)
                    dot_slots_ctl<double>(n, v64_.res.data(), v64_.res.data(), aux_set(1), done, s);
                else
                    dot_slots_ctl<V>(n, W.r.data(), W.r.data(), aux_set(1), done, s);
            }
        }
        else
            precond(true, k, pcur, done);
    }
    if(reliable)
    {
        // updates after the last fold (the loop ended without a stop)
        hipLaunchKernelGGL(k_fold<V>, dim3(grid_for(n, 256)), dim3(256), 0, s, n, x64, x,
                           (const int*)nullptr);
        SG_LAUNCH_CHECK();
    }
    int    hc[2];
    double hd[1];
    SG_HIP(hipMemcpyAsync(hc, ctl_.data(), sizeof(hc), hipMemcpyDeviceToHost, s));
    SG_HIP(hipMemcpyAsync(hd, ctl_d_.data(), sizeof(hd), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    if(!reliable)
        SG_CHECK(hc[1] == 0, -4, "PCG breakdown: non-finite r.z");
    st.breakdown  = hc[1];
    st.rz         = hd[0];
    st.converged  = hc[0] != 0 && hc[1] == 0;
    st.iterations = hc[0] != 0 ? hc[0] : (stalled_at ? stalled_at : maxit);
    st.stalled    = stalled_at != 0 && hc[0] == 0;
    return st;
}

// FP64 true relative residual ||b - A x|| / ||b|| (the FP64 operator)
double Pcg::true_residual(const SysView& S, const double* b, const double* x, hipStream_t s)
{
    const int n = 3 * S.A->nb;
    ensure(n);
    v64_.ensure(n);
    SG_HIP(hipMemsetAsync(aux_set(2), 0, 2 * kSlotSet * sizeof(double), s));
    residual(S, b, x, v64_.tmp.data(), s);
    dot_slots(n, v64_.tmp.data(), v64_.tmp.data(), aux_set(2), s);
    dot_slots(n, b, b, aux_set(3), s);
    double rr = read_slot_set(aux_set(2), s), bb = read_slot_set(aux_set(3), s);
    return bb > 0 ? sqrt(rr / bb) : 0.0;
}

PcgStats Pcg::solve(const SysView& S, Precond pc, Mas* mas, const double* b, double* x,
                    const PcgConfig& cfg, hipStream_t s)
{
    if(cfg.mixed)
    {
        SG_CHECK(!external_, -1, "a solver on a caller's workspace runs FP64 solves only");
        return solve_mixed(S, pc, mas, b, x, cfg, s);
    }
    PcgStats st          = solve_t<double>(S, pc, mas, b, x, cfg, s, nullptr, nullptr);
    st.outer_iterations  = 1;
    st.rel_residual      = true_residual(S, b, x, s);
    return st;
}

// Mixed precision: one PCG with the FP32 copy of the matrix, FP32 vectors
// and the FP32 MAS apply, in which the reference's every-10 residual
// recomputation is done in FP64 ("reliable updates"): at each check the FP32
// iterate is folded into an FP64 solution, r = b - A x is evaluated with the
// FP64 matrix and rounded back into the working residual, while the search
// direction is kept.  The stop test (r.r form) then measures the FP64 true
// residual.  Iterations stay close to the FP64 solve's; each costs FP32 traffic.
PcgStats Pcg::solve_mixed(const SysView& S, Precond pc, Mas* mas, const double* b, double* x,
                          const PcgConfig& cfg, hipStream_t s)
{
    const int n = 3 * S.A->nb;
    ensure(n);
    if(fp32_skip_ > 0)
    {
        // the recent FP32 attempts stalled: this solve runs in FP64
        fp32_skip_--;
        PcgStats f         = solve_t<double>(S, pc, mas, b, x, cfg, s, nullptr, nullptr);
        f.outer_iterations = 1;
        f.rel_residual     = true_residual(S, b, x, s);
        return f;
    }
    refresh_fp32(S, s);
    v64_.ensure(n);
    v32_.ensure(n);
    SG_HIP(hipMemsetAsync(x, 0, n * sizeof(double), s));
    convert<double, float>(n, b, v32_.rhs.data(), s);
    PcgConfig c32 = cfg;
    if(c32.max_iter <= 0)
        c32.max_iter = n < 20000 ? n : 20000;
    if(c32.stall_checks <= 0)
        c32.stall_checks = 20;
    PcgStats st = solve_t<float>(S, pc, mas, v32_.rhs.data(), v32_.sol.data(), c32, s, b, x);
    st.outer_iterations = 1;
    st.rel_residual     = true_residual(S, b, x, s);
    bool ok = st.breakdown == 0 && st.rel_residual == st.rel_residual
              && (st.converged || cfg.max_iter > 0);
    if(ok)
        fp32_pause_ = 0;
    else
    {
        fp32_pause_ = fp32_pause_ ? std::min(2 * fp32_pause_, 32) : 2;
        fp32_skip_  = fp32_pause_;
    }
    if(!ok)
    {
        // The FP32 iteration broke down, diverged or stalled (FP32 rounding
        // of a badly conditioned operator).  With a true-residual stop test
        // and an FP64 iterate better than zero, FP64 PCG continues from it
        // on A y = b - A x with the stop threshold of the original system;
        // otherwise FP64 solves again from zero.
        PcgStats f;
        bool     cont = !cfg.use_rz && st.breakdown != 1 && st.rel_residual == st.rel_residual
                    && st.rel_residual < 1.0;
        if(cont)
        {
            residual(S, b, x, v64_.res.data(), s);
            PcgConfig c64 = cfg;
            if(c64.q_ref <= 0.0)  // threshold of this system: |b|^2
            {
                SG_HIP(hipMemsetAsync(aux_set(3), 0, kSlotSet * sizeof(double), s));
                dot_slots(n, b, b, aux_set(3), s);
                c64.q_ref = read_slot_set(aux_set(3), s);
            }
            f = solve_t<double>(S, pc, mas, v64_.res.data(), v64_.sol.data(), c64, s, nullptr,
                                nullptr);
            launch_axpy_d(n, 1.0, v64_.sol.data(), x, s);
        }
        else
            f = solve_t<double>(S, pc, mas, b, x, cfg, s, nullptr, nullptr);
        f.iterations      += st.iterations;
        f.outer_iterations = 2;
        f.rel_residual     = true_residual(S, b, x, s);
        return f;
    }
    return st;
}

__global__ void k_axpy_d(int n, double a, const double* __restrict__ x, double* __restrict__ y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
        y[i] += a * x[i];
}

void launch_axpy_d(int n, double a, const double* x, double* y, hipStream_t s)
{
    hipLaunchKernelGGL(k_axpy_d, dim3(grid_for(n, 256)), dim3(256), 0, s, n, a, x, y);
    SG_LAUNCH_CHECK();
}

}  // namespace sgipc
