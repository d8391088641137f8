// linsys.hip -- BSR SpMV, residual, block-Jacobi and the PCG driver.
//
// PCG follows the reference's iteration exactly (pcg_solver.cu:171-304):
//   x = 0, r = b, z = M r, p = z;  for k = 1..: Ap = A p, a = rz/(p.Ap),
//   x += a p, r -= a Ap, (every 10th: r = b - A x), z = M r, rz' = r.z,
//   (every 10th: stop if rz' <= rel^2 rz0 or <= abs^2), p = z + (rz'/rz) p.
// Unlike the reference (two device->host copies per iteration, :145) every
// scalar stays on the device: pAp[k] and rz[k] live in a per-solve array and
// the kernels derive a and b from it; the host reads back only at the
// convergence test.
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

// ------------------------------------------------------------------ SpMV
// 8 lanes per block row; the row's blocks are contiguous so the 8 lanes of a
// group stream 576 contiguous bytes per step.
constexpr int kRowLanes = 8;

template <bool RESID>
__global__ void __launch_bounds__(256) k_spmv_csr(int nb, const int* __restrict__ rowptr,
                                                  const int* __restrict__ cols,
                                                  const double* __restrict__ vals,
                                                  const double* __restrict__ x,
                                                  const double* __restrict__ b,
                                                  double* __restrict__ y, double* dot)
{
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
            const double* B  = vals + 9 * (size_t)k;
            const double* xj = x + 3 * (size_t)cols[k];
            double        x0 = xj[0], x1 = xj[1], x2 = xj[2];
            a0 += B[0] * x0 + B[1] * x1 + B[2] * x2;
            a1 += B[3] * x0 + B[4] * x1 + B[5] * x2;
            a2 += B[6] * x0 + B[7] * x1 + B[8] * x2;
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
        if(RESID)
        {
            y[3 * row + 0] = b[3 * row + 0] - a0;
            y[3 * row + 1] = b[3 * row + 1] - a1;
            y[3 * row + 2] = b[3 * row + 2] - a2;
        }
        else
        {
            y[3 * row + 0] = a0;
            y[3 * row + 1] = a1;
            y[3 * row + 2] = a2;
            if(dot)
                d = x[3 * row + 0] * a0 + x[3 * row + 1] * a1 + x[3 * row + 2] * a2;
        }
    }
    if(!RESID && dot)
    {
        d = wave_sum(d);
        if(__lane_id() == 0 && d != 0.0)
            atomicAdd(dot, d);
    }
}

// COO tail: y[r] += B x[c]  (sign = -1 for residuals)
__global__ void k_spmv_coo(const unsigned* __restrict__ count, unsigned cap,
                           const int* __restrict__ r, const int* __restrict__ c,
                           const double* __restrict__ v, const double* __restrict__ x, double* y,
                           double* dot, double sign)
{
    unsigned n = *count;
    n          = n < cap ? n : cap;
    unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    double   d = 0.0;
    if(t < n)
    {
        const double* B  = v + 9 * (size_t)t;
        const double* xj = x + 3 * (size_t)c[t];
        int           i  = r[t];
        double        a0 = B[0] * xj[0] + B[1] * xj[1] + B[2] * xj[2];
        double        a1 = B[3] * xj[0] + B[4] * xj[1] + B[5] * xj[2];
        double        a2 = B[6] * xj[0] + B[7] * xj[1] + B[8] * xj[2];
        atomicAdd(y + 3 * i + 0, sign * a0);
        atomicAdd(y + 3 * i + 1, sign * a1);
        atomicAdd(y + 3 * i + 2, sign * a2);
        if(dot)
            d = x[3 * i + 0] * a0 + x[3 * i + 1] * a1 + x[3 * i + 2] * a2;
    }
    if(dot)
    {
        d = wave_sum(d);
        if(__lane_id() == 0 && d != 0.0)
            atomicAdd(dot, d);
    }
}

static void launch_coo(const Coo* coo, const double* x, double* y, double* dot, double sign,
                       hipStream_t s)
{
    if(!coo || coo->cap == 0)
        return;
    // the live count is device-resident; launch over the capacity
    hipLaunchKernelGGL(k_spmv_coo, dim3(grid_for(coo->cap, 256)), dim3(256), 0, s,
                       coo->count.data(), (unsigned)coo->cap, coo->r.data(), coo->c.data(),
                       coo->v.data(), x, y, dot, sign);
    SG_LAUNCH_CHECK();
}

void spmv(const SysView& S, const double* x, double* y, double* dot, hipStream_t s)
{
    const Bsr& A = *S.A;
    hipLaunchKernelGGL(k_spmv_csr<false>, dim3(grid_for((long long)A.nb * kRowLanes, 256)),
                       dim3(256), 0, s, A.nb, A.rowptr.data(), A.cols.data(), A.vals.data(), x,
                       (const double*)nullptr, y, dot);
    SG_LAUNCH_CHECK();
    launch_coo(S.coo, x, y, dot, 1.0, s);
}

void residual(const SysView& S, const double* b, const double* x, double* r, hipStream_t s)
{
    const Bsr& A = *S.A;
    hipLaunchKernelGGL(k_spmv_csr<true>, dim3(grid_for((long long)A.nb * kRowLanes, 256)),
                       dim3(256), 0, s, A.nb, A.rowptr.data(), A.cols.data(), A.vals.data(), x,
                       b, r, (double*)nullptr);
    SG_LAUNCH_CHECK();
    launch_coo(S.coo, x, r, nullptr, -1.0, s);
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

// optional axpy (x += a p, r -= a Ap with a = rz_prev / pAp), then z = Dinv r
// and rz += r.z
__global__ void k_update_diag(int nb, double* x, double* r, const double* p, const double* Ap,
                              const double* rz_prev, const double* pAp,
                              const double* __restrict__ dinv, double* z, double* rz_out)
{
    int    i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.0;
    if(i < nb)
    {
        double ri[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
        if(x)
        {
            double a = *rz_prev / *pAp;
            for(int c = 0; c < 3; c++)
            {
                x[3 * i + c] += a * p[3 * i + c];
                ri[c] -= a * Ap[3 * i + c];
                r[3 * i + c] = ri[c];
            }
        }
        if(z)
        {
            const double* D = dinv + 9 * (size_t)i;
            for(int c = 0; c < 3; c++)
            {
                double zc    = D[3 * c] * ri[0] + D[3 * c + 1] * ri[1] + D[3 * c + 2] * ri[2];
                z[3 * i + c] = zc;
                d += ri[c] * zc;
            }
        }
    }
    if(rz_out)
    {
        d = wave_sum(d);
        if(__lane_id() == 0 && d != 0.0)
            atomicAdd(rz_out, d);
    }
}

__global__ void k_axpy(int n, double* x, double* r, const double* p, const double* Ap,
                       const double* rz_prev, const double* pAp)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n)
        return;
    double a = *rz_prev / *pAp;
    x[i] += a * p[i];
    r[i] -= a * Ap[i];
}

__global__ void k_pupdate(int n, double* p, const double* z, const double* rz_new,
                          const double* rz_old)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n)
        return;
    double b = *rz_new / *rz_old;
    p[i]     = z[i] + b * p[i];
}

__global__ void k_dot(int n, const double* a, const double* b, double* out)
{
    int    i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = (i < n) ? a[i] * b[i] : 0.0;
    d        = wave_sum(d);
    if(__lane_id() == 0 && d != 0.0)
        atomicAdd(out, d);
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
                       dinv_.data(), z, (double*)nullptr);
    SG_LAUNCH_CHECK();
}

void Pcg::ensure(int n, int maxit)
{
    if(n != n_)
    {
        r_.resize(n);
        z_.resize(n);
        p_.resize(n);
        Ap_.resize(n);
        tmp_.resize(n);
        n_ = n;
    }
    if(maxit + 2 > maxit_)
    {
        sc_.resize(2 * (size_t)(maxit + 2) + 8);
        maxit_ = maxit + 2;
    }
}

static double read_scalar(const double* d, hipStream_t s)
{
    double h = 0;
    SG_HIP(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    return h;
}

PcgStats Pcg::solve(const SysView& S, Precond pc, Mas* mas, const double* b, double* x,
                    const PcgConfig& cfg, hipStream_t s)
{
    const int nb    = S.A->nb;
    const int n     = 3 * nb;
    int       maxit = cfg.max_iter > 0 ? cfg.max_iter : n;
    int       every = cfg.check_every > 0 ? cfg.check_every : 10;
    ensure(n, maxit);
    double* rz  = sc_.data();                 // rz[k], k = 0..maxit
    double* pAp = sc_.data() + (maxit + 2);   // pAp[k]
    double* aux = sc_.data() + 2 * (maxit + 2); // rr0, rr, |r|^2, |b|^2
    SG_HIP(hipMemsetAsync(sc_.data(), 0, sc_.size() * sizeof(double), s));
    SG_HIP(hipMemsetAsync(x, 0, n * sizeof(double), s));
    SG_HIP(hipMemcpyAsync(r_.data(), b, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    if(pc == Precond::Diag)
        build_diag(*S.A, s);

    auto precond = [&](bool axpy, int k) {
        if(pc == Precond::Mas)
            mas->apply_fused(axpy ? x : nullptr, r_.data(), p_.data(), Ap_.data(),
                             axpy ? rz + (k - 1) : nullptr, axpy ? pAp + k : nullptr, z_.data(),
                             rz + k, s);
        else
        {
            hipLaunchKernelGGL(k_update_diag, dim3(grid_for(nb, 256)), dim3(256), 0, s, nb,
                               axpy ? x : nullptr, r_.data(), p_.data(), Ap_.data(),
                               axpy ? rz + (k - 1) : nullptr, axpy ? pAp + k : nullptr,
                               dinv_.data(), z_.data(), rz + k);
            SG_LAUNCH_CHECK();
        }
    };
    auto dot = [&](const double* a, const double* bb, double* out) {
        hipLaunchKernelGGL(k_dot, dim3(grid_for(n, 256)), dim3(256), 0, s, n, a, bb, out);
        SG_LAUNCH_CHECK();
    };

    PcgStats st;
    precond(false, 0);
    SG_HIP(hipMemcpyAsync(p_.data(), z_.data(), n * sizeof(double), hipMemcpyDeviceToDevice, s));
    if(!cfg.use_rz)
        dot(r_.data(), r_.data(), aux + 0);
    double rz0 = read_scalar(rz, s);
    double rr0 = cfg.use_rz ? 0.0 : read_scalar(aux + 0, s);
    st.rz0     = rz0;
    int k      = 0;
    if(rz0 == 0.0 || !(rz0 == rz0))
    {
        st.iterations = 0;
        st.converged  = (rz0 == 0.0);
        return st;
    }
    double rel2 = cfg.rel_tol * cfg.rel_tol, abs2 = cfg.abs_tol * cfg.abs_tol;
    for(k = 1; k < maxit; k++)
    {
        if(prof)
            prof->begin("spmv", s);
        spmv(S, p_.data(), Ap_.data(), pAp + k, s);
        if(prof)
            prof->end("spmv", s);
        if(k % every == 0)
        {
            hipLaunchKernelGGL(k_axpy, dim3(grid_for(n, 256)), dim3(256), 0, s, n, x, r_.data(),
                               p_.data(), Ap_.data(), rz + (k - 1), pAp + k);
            SG_LAUNCH_CHECK();
            residual(S, b, x, r_.data(), s);
            precond(false, k);
            double rzk = read_scalar(rz + k, s);
            st.rz      = rzk;
            bool stop;
            if(cfg.use_rz)
                stop = (rzk <= rel2 * rz0 || rzk <= abs2);
            else
            {
                SG_HIP(hipMemsetAsync(aux + 1, 0, sizeof(double), s));
                dot(r_.data(), r_.data(), aux + 1);
                double rr = read_scalar(aux + 1, s);
                stop      = (rr <= rel2 * rr0 || rr <= abs2);
            }
            SG_CHECK(rzk == rzk, -4, "PCG breakdown: non-finite r.z");
            if(stop)
            {
                st.converged = true;
                break;
            }
        }
        else
            precond(true, k);
        hipLaunchKernelGGL(k_pupdate, dim3(grid_for(n, 256)), dim3(256), 0, s, n, p_.data(),
                           z_.data(), rz + k, rz + (k - 1));
        SG_LAUNCH_CHECK();
    }
    st.iterations = k;
    // FP64 true relative residual
    residual(S, b, x, tmp_.data(), s);
    dot(tmp_.data(), tmp_.data(), aux + 2);
    dot(b, b, aux + 3);
    double rr = read_scalar(aux + 2, s), bb = read_scalar(aux + 3, s);
    st.rel_residual = bb > 0 ? sqrt(rr / bb) : 0.0;
    return st;
}

}  // namespace sgipc
