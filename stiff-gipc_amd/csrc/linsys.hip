// linsys.hip -- BSR SpMV, residual, block-Jacobi and the PCG driver.
//
// PCG follows the reference's iteration exactly (pcg_solver.cu:171-304):
//   x = 0, r = b, z = M r, p = z;  for k = 1..: Ap = A p, a = rz/(p.Ap),
//   x += a p, r -= a Ap, (every 10th: r = b - A x), z = M r, rz' = r.z,
//   (every 10th: stop if rz' <= rel^2 rz0 or <= abs^2), p = z + (rz'/rz) p.
// Unlike the reference (two device->host copies per iteration, :145) every
// scalar stays on the device: pAp[k] and rz[k] are accumulated into slot sets
// (see common.h) and the consuming kernels derive a and b from them; the
// host reads back only at the convergence test.
#include "linsys.h"

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

// exclusive scan of a[0..n) into out[0..n], out[n] = total (one workgroup of
// 1024 threads, each owning a contiguous chunk)
__global__ void __launch_bounds__(1024) k_scan_one_block(int n, const int* __restrict__ a,
                                                         int* __restrict__ out)
{
    __shared__ int part[1024];
    const int t     = threadIdx.x;
    const int chunk = (n + 1023) / 1024;
    const int b0    = t * chunk;
    const int b1    = (b0 + chunk < n) ? b0 + chunk : n;
    int       sum   = 0;
    for(int i = b0; i < b1; i++)
        sum += a[i];
    part[t] = sum;
    __syncthreads();
    for(int o = 1; o < 1024; o <<= 1)
    {
        int v = (t >= o) ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - sum;  // exclusive prefix of this chunk
    for(int i = b0; i < b1; i++)
    {
        out[i] = run;
        run += a[i];
    }
    if(t == 1023)
        out[n] = part[1023];
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

void Coo::build_rows(hipStream_t s)
{
    SG_CHECK(nb > 0, -5, "contact CSR: reserve() not called");
    SG_HIP(hipMemsetAsync(fill.data(), 0, (nb + 1) * sizeof(int), s));
    hipLaunchKernelGGL(k_coo_hist, dim3(grid_for(cap, 256)), dim3(256), 0, s, count.data(),
                       (unsigned)cap, r.data(), fill.data());
    SG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_one_block, dim3(1), dim3(1024), 0, s, nb, fill.data(), rp.data());
    SG_LAUNCH_CHECK();
    SG_HIP(hipMemsetAsync(fill.data(), 0, (nb + 1) * sizeof(int), s));
    hipLaunchKernelGGL(k_coo_scatter, dim3(grid_for(cap, 256)), dim3(256), 0, s, count.data(),
                       (unsigned)cap, r.data(), c.data(), v.data(), rp.data(), fill.data(),
                       rc.data(), rv.data());
    SG_LAUNCH_CHECK();
    rows_valid = true;
}

// ------------------------------------------------------------------ SpMV
// 8 lanes per block row; a row's blocks are contiguous, so the 8 lanes of a
// group stream 576 contiguous bytes per step.  The static pattern and the
// row-sorted contact blocks of the row are traversed by the same lanes, so y
// is written once per row without atomics.
//
// MODE_PLAIN: y = A x            (+ optional x.y into a slot set)
// MODE_RESID: y = b - A x
// MODE_PCG:   the PCG direction is formed on the fly: for every column j read,
//             p_j = z_j + beta p_prev_j (beta = rz[k-1] / rz[k-2], slot sets),
//             or p_j = z_j on the first iteration; the row owner stores p_i,
//             y = A p and p.y is accumulated into a slot set.  Every read of
//             p_j evaluates the same fma, so all rows see identical values.
constexpr int kRowLanes = 8;
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
    int           check_k  = 0;
};

template <int MODE>
__global__ void __launch_bounds__(256)
    k_spmv(int nb, const int* __restrict__ rowptr, const int* __restrict__ cols,
           const double* __restrict__ vals, const int* __restrict__ rp2,
           const int* __restrict__ cols2, const double* __restrict__ vals2,
           const double* __restrict__ x, const double* __restrict__ pprev,
           double* __restrict__ pcur, const double* __restrict__ b, double* __restrict__ y,
           const double* rz_a, const double* rz_b, double* dot, PcgCtl pc)
{
    if(pc.ctl && pc.ctl[0])
        return;
    if(MODE == MODE_PCG && pc.q_set)
    {
        double q   = slot_set_sum(pc.q_set);
        double rzc = slot_set_sum(pc.rz_check);
        bool   bad = !(rzc == rzc);
        if(blockIdx.x == 0 && threadIdx.x == 0)
            pc.ctl_d[0] = rzc;
        if(q <= pc.thr || bad)
        {
            if(blockIdx.x == 0 && threadIdx.x == 0)
            {
                pc.ctl[1] = bad ? 1 : 0;
                pc.ctl[0] = pc.check_k;
            }
            return;
        }
    }
    double beta = 0.0;
    if(MODE == MODE_PCG && pprev)
        beta = slot_set_sum(rz_a) / slot_set_sum(rz_b);
    auto fetch = [&](int j, double& x0, double& x1, double& x2) {
        const double* xj = x + 3 * (size_t)j;
        if(MODE == MODE_PCG && pprev)
        {
            const double* pj = pprev + 3 * (size_t)j;
            x0               = fma(beta, pj[0], xj[0]);
            x1               = fma(beta, pj[1], xj[1]);
            x2               = fma(beta, pj[2], xj[2]);
        }
        else
        {
            x0 = xj[0];
            x1 = xj[1];
            x2 = xj[2];
        }
    };
    int  gid  = blockIdx.x * blockDim.x + threadIdx.x;
    int  row  = gid / kRowLanes;
    int  g    = gid % kRowLanes;
    bool live = row < nb;
    double a0 = 0, a1 = 0, a2 = 0;
    if(live)
    {
        int k1 = rowptr[row + 1];
        for(int k = rowptr[row] + g; k < k1; k += kRowLanes)
        {
            const double* B = vals + 9 * (size_t)k;
            double        x0, x1, x2;
            fetch(cols[k], x0, x1, x2);
            a0 += B[0] * x0 + B[1] * x1 + B[2] * x2;
            a1 += B[3] * x0 + B[4] * x1 + B[5] * x2;
            a2 += B[6] * x0 + B[7] * x1 + B[8] * x2;
        }
        if(rp2)
        {
            int q1 = rp2[row + 1];
            for(int k = rp2[row] + g; k < q1; k += kRowLanes)
            {
                const double* B = vals2 + 9 * (size_t)k;
                double        x0, x1, x2;
                fetch(cols2[k], x0, x1, x2);
                a0 += B[0] * x0 + B[1] * x1 + B[2] * x2;
                a1 += B[3] * x0 + B[4] * x1 + B[5] * x2;
                a2 += B[6] * x0 + B[7] * x1 + B[8] * x2;
            }
        }
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
            if(MODE == MODE_PCG || dot)
            {
                double x0, x1, x2;
                fetch(row, x0, x1, x2);
                if(MODE == MODE_PCG)
                {
                    pcur[i + 0] = x0;
                    pcur[i + 1] = x1;
                    pcur[i + 2] = x2;
                }
                d = x0 * a0 + x1 * a1 + x2 * a2;
            }
        }
    }
    if(MODE == MODE_PCG || (MODE == MODE_PLAIN && dot))
        slot_set_add(dot, d);
}

static inline void contact_rows(const SysView& S, const int*& rp2, const int*& c2,
                                const double*& v2)
{
    rp2 = nullptr;
    c2  = nullptr;
    v2  = nullptr;
    if(S.coo && S.coo->cap > 0)
    {
        SG_CHECK(S.coo->rows_valid, -5, "contact CSR is stale: build_rows() after assembly");
        rp2 = S.coo->rp.data();
        c2  = S.coo->rc.data();
        v2  = S.coo->rv.data();
    }
}

void spmv(const SysView& S, const double* x, double* y, double* dot, hipStream_t s)
{
    const Bsr&    A = *S.A;
    const int*    rp2;
    const int*    c2;
    const double* v2;
    contact_rows(S, rp2, c2, v2);
    hipLaunchKernelGGL(k_spmv<MODE_PLAIN>, dim3(grid_for((long long)A.nb * kRowLanes, 256)),
                       dim3(256), 0, s, A.nb, A.rowptr.data(), A.cols.data(), A.vals.data(), rp2,
                       c2, v2, x, (const double*)nullptr, (double*)nullptr,
                       (const double*)nullptr, y, (const double*)nullptr,
                       (const double*)nullptr, dot, PcgCtl{});
    SG_LAUNCH_CHECK();
}

static void residual_ctl(const SysView& S, const double* b, const double* x, double* r,
                        const PcgCtl& pc, hipStream_t s)
{
    const Bsr&    A = *S.A;
    const int*    rp2;
    const int*    c2;
    const double* v2;
    contact_rows(S, rp2, c2, v2);
    hipLaunchKernelGGL(k_spmv<MODE_RESID>, dim3(grid_for((long long)A.nb * kRowLanes, 256)),
                       dim3(256), 0, s, A.nb, A.rowptr.data(), A.cols.data(), A.vals.data(), rp2,
                       c2, v2, x, (const double*)nullptr, (double*)nullptr, b, r,
                       (const double*)nullptr, (const double*)nullptr, (double*)nullptr, pc);
    SG_LAUNCH_CHECK();
}

void residual(const SysView& S, const double* b, const double* x, double* r, hipStream_t s)
{
    residual_ctl(S, b, x, r, PcgCtl{}, s);
}

// PCG: p = z (+ beta p_prev), Ap = A p, p.Ap -> slot set
static void spmv_pcg(const SysView& S, const double* z, const double* pprev, double* pcur,
                     double* Ap, const double* rz_a, const double* rz_b, double* pap,
                     const PcgCtl& pc, hipStream_t s)
{
    const Bsr&    A = *S.A;
    const int*    rp2;
    const int*    c2;
    const double* v2;
    contact_rows(S, rp2, c2, v2);
    hipLaunchKernelGGL(k_spmv<MODE_PCG>, dim3(grid_for((long long)A.nb * kRowLanes, 256)),
                       dim3(256), 0, s, A.nb, A.rowptr.data(), A.cols.data(), A.vals.data(), rp2,
                       c2, v2, z, pprev, pcur, (const double*)nullptr, Ap, rz_a, rz_b, pap, pc);
    SG_LAUNCH_CHECK();
}

__global__ void __launch_bounds__(256) k_dot_slots(int n, const double* __restrict__ a,
                                                   const double* __restrict__ b, double* out,
                                                   const int* __restrict__ done)
{
    if(done && *done)
        return;
    int    i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.0;
    for(; i < n; i += gridDim.x * blockDim.x)
        d += a[i] * b[i];
    slot_set_add(out, d);
}

static void dot_slots_ctl(int n, const double* a, const double* b, double* out,
                          const int* done, hipStream_t s)
{
    unsigned g = grid_for(n, 256);
    g          = g > 1024 ? 1024 : g;
    hipLaunchKernelGGL(k_dot_slots, dim3(g), dim3(256), 0, s, n, a, b, out, done);
    SG_LAUNCH_CHECK();
}

void dot_slots(int n, const double* a, const double* b, double* out, hipStream_t s)
{
    dot_slots_ctl(n, a, b, out, nullptr, s);
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
__global__ void __launch_bounds__(256)
    k_update_diag(int nb, double* x, double* r, const double* p, const double* Ap,
                  const double* rz_prev, const double* pap, const double* __restrict__ dinv,
                  double* z, double* rz_out, double* zero0, double* zero1,
                  const int* __restrict__ done)
{
    if(done && *done)
        return;
    double a = 0.0;
    if(x)
        a = slot_set_sum(rz_prev) / slot_set_sum(pap);
    int    i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.0;
    if(i < nb)
    {
        double ri[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
        if(x)
        {
            for(int c = 0; c < 3; c++)
            {
                x[3 * i + c] += a * p[3 * i + c];
                ri[c] -= a * Ap[3 * i + c];
                r[3 * i + c] = ri[c];
            }
        }
        const double* D = dinv + 9 * (size_t)i;
        for(int c = 0; c < 3; c++)
        {
            double zc    = D[3 * c] * ri[0] + D[3 * c + 1] * ri[1] + D[3 * c + 2] * ri[2];
            z[3 * i + c] = zc;
            d += ri[c] * zc;
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

__global__ void __launch_bounds__(256) k_axpy(int n, double* x, double* r, const double* p,
                                              const double* Ap, const double* rz_prev,
                                              const double* pap, const int* __restrict__ done)
{
    if(done && *done)
        return;
    double a = slot_set_sum(rz_prev) / slot_set_sum(pap);
    int    i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n)
    {
        x[i] += a * p[i];
        r[i] -= a * Ap[i];
    }
}

void Pcg::build_diag(const Bsr& A, hipStream_t s)
{
    dinv_.resize((size_t)9 * A.nb);
    hipLaunchKernelGGL(k_diag_inverse, dim3(grid_for(A.nb, 256)), dim3(256), 0, s, A.nb,
                       A.diag.data(), A.vals.data(), dinv_.data());
    SG_LAUNCH_CHECK();
}

void Pcg::apply_diag(int nb, const double* r, double* z, hipStream_t s)
{
    hipLaunchKernelGGL(k_update_diag, dim3(grid_for(nb, 256)), dim3(256), 0, s, nb,
                       (double*)nullptr, const_cast<double*>(r), (const double*)nullptr,
                       (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                       dinv_.data(), z, (double*)nullptr, (double*)nullptr, (double*)nullptr,
                       (const int*)nullptr);
    SG_LAUNCH_CHECK();
}


Pcg::~Pcg()
{
    if(h_ctl_)
        (void)hipHostFree(h_ctl_);
    for(auto& e : ev_)
        if(e)
            (void)hipEventDestroy(e);
}

void Pcg::ensure(int n)
{
    if(n != n_)
    {
        r_.resize(n);
        z_.resize(n);
        p_[0].resize(n);
        p_[1].resize(n);
        Ap_.resize(n);
        tmp_.resize(n);
        n_ = n;
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
// of iteration k.
PcgStats Pcg::solve(const SysView& S, Precond pc, Mas* mas, const double* b, double* x,
                    const PcgConfig& cfg, hipStream_t s)
{
    const int nb    = S.A->nb;
    const int n     = 3 * nb;
    int       maxit = cfg.max_iter > 0 ? cfg.max_iter : n;
    int       every = cfg.check_every > 0 ? cfg.check_every : 10;
    ensure(n);
    SG_HIP(hipMemsetAsync(red_.data(), 0, red_.size() * sizeof(double), s));
    SG_HIP(hipMemsetAsync(ctl_.data(), 0, ctl_.size() * sizeof(int), s));
    SG_HIP(hipMemsetAsync(ctl_d_.data(), 0, ctl_d_.size() * sizeof(double), s));
    SG_HIP(hipMemsetAsync(x, 0, n * sizeof(double), s));
    SG_HIP(hipMemcpyAsync(r_.data(), b, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    if(pc == Precond::Diag)
        build_diag(*S.A, s);
    const int* done = ctl_.data();

    // z = M r with the optional update; rz[k] -> rz_set(k)
    auto precond = [&](bool axpy, int k, const double* p, const int* dn) {
        double* rz_prev = axpy ? rz_set(k - 1) : nullptr;
        double* pap     = axpy ? pap_set(k) : nullptr;
        double* zero0   = pap_set(k + 2);
        double* zero1   = rz_set(k + 2);
        if(pc == Precond::Mas)
            mas->apply_fused(axpy ? x : nullptr, r_.data(), p, Ap_.data(), rz_prev, pap,
                             z_.data(), rz_set(k), zero0, zero1, dn, s);
        else
        {
            hipLaunchKernelGGL(k_update_diag, dim3(grid_for(nb, 256)), dim3(256), 0, s, nb,
                               axpy ? x : nullptr, r_.data(), p, Ap_.data(), rz_prev, pap,
                               dinv_.data(), z_.data(), rz_set(k), zero0, zero1, dn);
            SG_LAUNCH_CHECK();
        }
    };

    PcgStats st;
    precond(false, 0, nullptr, nullptr);
    if(!cfg.use_rz)
        dot_slots(n, r_.data(), r_.data(), aux_set(0), s);
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
    double q0   = cfg.use_rz ? rz0 : rr0;
    double thr  = rel2 * q0 > abs2 ? rel2 * q0 : abs2;

    PcgCtl ctl_only;
    ctl_only.ctl   = ctl_.data();
    ctl_only.ctl_d = ctl_d_.data();
    int  slot = 0, pend_slot = 0;
    bool have_pending = false;
    for(int k = 1; k <= maxit; k++)
    {
        PcgCtl     pcx  = ctl_only;
        const bool test = (k - 1 >= 1) && ((k - 1) % every == 0);
        if(test)
        {
            pcx.q_set    = cfg.use_rz ? rz_set(k - 1) : aux_set(1);
            pcx.rz_check = rz_set(k - 1);
            pcx.thr      = thr;
            pcx.check_k  = k - 1;
        }
        double*       pcur  = p_[k & 1].data();
        const double* pprev = (k == 1) ? nullptr : p_[(k - 1) & 1].data();
        if(prof)
            prof->begin("spmv", s);
        spmv_pcg(S, z_.data(), pprev, pcur, Ap_.data(), rz_set(k - 1),
                 k >= 2 ? rz_set(k - 2) : rz_set(k - 1), pap_set(k), pcx, s);
        if(prof)
            prof->end("spmv", s);
        if(test)
        {
            SG_HIP(hipMemcpyAsync(h_ctl_ + 4 * slot, ctl_.data(), 2 * sizeof(int),
                                  hipMemcpyDeviceToHost, s));
            SG_HIP(hipEventRecord(ev_[slot], s));
            if(have_pending)
            {
                SG_HIP(hipEventSynchronize(ev_[pend_slot]));
                if(h_ctl_[4 * pend_slot] != 0)
                    break;
            }
            have_pending = true;
            pend_slot    = slot;
            slot ^= 1;
        }
        if(k == maxit)
            break;  // the last launch only evaluated the final stop test
        if(k % every == 0)
        {
            hipLaunchKernelGGL(k_axpy, dim3(grid_for(n, 256)), dim3(256), 0, s, n, x, r_.data(),
                               pcur, Ap_.data(), rz_set(k - 1), pap_set(k), done);
            SG_LAUNCH_CHECK();
            residual_ctl(S, b, x, r_.data(), ctl_only, s);
            precond(false, k, pcur, done);
            if(!cfg.use_rz)
            {
                hipLaunchKernelGGL(k_clear_set, dim3(1), dim3(64), 0, s, aux_set(1), done);
                SG_LAUNCH_CHECK();
                dot_slots_ctl(n, r_.data(), r_.data(), aux_set(1), done, s);
            }
        }
        else
            precond(true, k, pcur, done);
    }
    int    hc[2];
    double hd[1];
    SG_HIP(hipMemcpyAsync(hc, ctl_.data(), sizeof(hc), hipMemcpyDeviceToHost, s));
    SG_HIP(hipMemcpyAsync(hd, ctl_d_.data(), sizeof(hd), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    SG_CHECK(hc[1] == 0, -4, "PCG breakdown: non-finite r.z");
    st.rz         = hd[0];
    st.converged  = hc[0] != 0;
    st.iterations = hc[0] != 0 ? hc[0] : maxit;
    // FP64 true relative residual
    residual(S, b, x, tmp_.data(), s);
    dot_slots(n, tmp_.data(), tmp_.data(), aux_set(2), s);
    dot_slots(n, b, b, aux_set(3), s);
    double rr = read_slot_set(aux_set(2), s), bb = read_slot_set(aux_set(3), s);
    st.rel_residual = bb > 0 ? sqrt(rr / bb) : 0.0;
    return st;
}

}  // namespace sgipc
