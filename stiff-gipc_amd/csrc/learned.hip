// learned.hip -- learned affine coarse warm start for the MAS-PCG solve.
//
// No counterpart in the reference.  MAS's coarse spaces are piecewise
// constant (one translation per aggregate), so the rotation / shear modes
// of stiff bodies are left to the Krylov iteration.  Each aggregate c of the
// coarsest MAS level (its free FEM nodes) gets an affine coarse space
// x_i = p + A (x_i - centroid) -- J_i q, q in R^12, the affine-body map -- and
//   q_c = (P_c^T A P_c)^-1 P_c^T b           (block-diagonal Galerkin solve)
//   q_c += s_c MLP(q_c / s_c, r_c / |r_c|)   (learned correction, s_c = |q_c|,
//                                             r_c = P_c^T b; 24 -> 128 -> 128 -> 12,
//                                             ReLU, BF16 weights, FP32 accumulation)
// The network corrects for the coupling between aggregates (contacts, shared
// banks) that the block-diagonal solve leaves out.  The PCG solve continues
// from x0 = P q (A y = b - A x0) with the stop threshold of the zero-guess
// solve; plain MAS judges the guess (kept only when it lowers the
// preconditioned residual) and the zero-guess solve is the fallback.
//
// Training (full batch, Adam on FP32 master weights, BF16 copies in the
// forward pass): targets are (q*_c - q_c) / s_c with q*_c the least-squares
// affine fit of the converged Newton direction on the aggregate.
#include <stdint.h>

#include <algorithm>

#include "../../include/sgipc.h"
#include "learned.h"

namespace sgipc
{

__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f_to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

// ------------------------------------------------------------- aggregates
// aggregate of every node: free FEM nodes of an aggregate of level lev
__global__ void k_agg_of(int nn, int nfem, const int* __restrict__ btype, int L, int lev,
                         const int* __restrict__ chain, int off, int nc, int* __restrict__ agg,
                         unsigned* __restrict__ cnt)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    int c = -1;
    if(v < nfem && btype[v] == 0)
    {
        c = chain[(size_t)v * (L - 1) + lev - 1] - off;
        if(c < 0 || c >= nc)
            c = -1;
    }
    if(v < nn)
        agg[v] = c;
    // few coarsest aggregates: per-node atomics on them would serialise
    wave_append_keyed(cnt, c);
}

__global__ void k_agg_fill(int nn, const int* __restrict__ agg, const int* __restrict__ ptr,
                           unsigned* __restrict__ cur, int* __restrict__ nodes)
{
    int      v    = blockIdx.x * blockDim.x + threadIdx.x;
    int      c    = v < nn ? agg[v] : -1;
    unsigned slot = wave_append_keyed(cur, c);
    if(c >= 0)
        nodes[ptr[c] + (int)slot] = v;
}

// one 64-lane workgroup per aggregate: centroid of its nodes
__global__ void __launch_bounds__(64) k_agg_centroid(const int* __restrict__ ptr,
                                                     const int* __restrict__ nodes,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ cen)
{
    int    c = blockIdx.x;
    double s[3] = {0, 0, 0};
    for(int k = ptr[c] + (int)threadIdx.x; k < ptr[c + 1]; k += 64)
    {
        int v = nodes[k];
#pragma unroll
        for(int e = 0; e < 3; e++)
            s[e] += x[3 * (size_t)v + e];
    }
    int n = ptr[c + 1] - ptr[c];
#pragma unroll
    for(int e = 0; e < 3; e++)
    {
        double t = wave_sum(s[e]);
        if(threadIdx.x == 0)
            cen[3 * c + e] = n ? t / n : 0.0;
    }
}

// upper (P <= Q) index of a symmetric 12x12
__device__ __forceinline__ int u12(int p, int q) { return p * 12 - p * (p - 1) / 2 + (q - p); }

// accumulate C_m(xi) A C_n(xj)^T into the 78 upper entries of the 12x12
__device__ __forceinline__ void acc_jaj(double (&U)[78], const double* A, const double xi[3],
                                        const double xj[3])
{
#pragma unroll
    for(int m = 0; m < 4; m++)
#pragma unroll
        for(int n = 0; n < 4; n++)
#pragma unroll
            for(int a = 0; a < 3; a++)
#pragma unroll
                for(int b = 0; b < 3; b++)
                {
                    int P = 3 * m + a, Q = 3 * n + b;
                    if(P > Q)
                        continue;
                    // (C_m A)[a][c'] then times C_n^T
                    double v;
                    if(n == 0)
                        v = (m == 0) ? A[3 * a + b] : xi[a] * A[3 * (m - 1) + b];
                    else
                        v = ((m == 0) ? A[3 * a + (n - 1)] : xi[a] * A[3 * (m - 1) + (n - 1)]) * xj[b];
                    U[u12(P, Q)] += v;
                }
}

// Galerkin matrix P^T A P (upper 78) and P^T b over the blocks of the static
// pattern and the contact rows whose column node is in the same aggregate.
// G single-wave workgroups per aggregate (grid nc * G), one node per lane per
// pass; each workgroup's 90 partial sums are reduced through LDS (one row
// per lane, fixed order) and added into Ac / rc, which start at zero.
__global__ void __launch_bounds__(64) k_agg_galerkin(
    int G, const int* __restrict__ ptr, const int* __restrict__ nodes, const int* __restrict__ agg,
    const double* __restrict__ x, const double* __restrict__ cen,
    const int* __restrict__ rowptr, const int* __restrict__ cols, const double* __restrict__ vals,
    const int* __restrict__ crp, const int* __restrict__ crc, const double* __restrict__ crv,
    const double* __restrict__ b, double* __restrict__ Ac, double* __restrict__ rc)
{
    __shared__ double red[90][65];
    const int c = blockIdx.x / G, gpart = blockIdx.x % G;
    double    U[78], r[12];
#pragma unroll
    for(int k = 0; k < 78; k++)
        U[k] = 0.0;
#pragma unroll
    for(int k = 0; k < 12; k++)
        r[k] = 0.0;
    double cx = cen[3 * c], cy = cen[3 * c + 1], cz = cen[3 * c + 2];
    for(int k = ptr[c] + gpart * 64 + (int)threadIdx.x; k < ptr[c + 1]; k += 64 * G)
    {
        int    i     = nodes[k];
        double xi[3] = {x[3 * (size_t)i] - cx, x[3 * (size_t)i + 1] - cy, x[3 * (size_t)i + 2] - cz};
        // r += J_i^T b_i
#pragma unroll
        for(int e = 0; e < 3; e++)
        {
            double be = b[3 * (size_t)i + e];
            r[e] += be;
#pragma unroll
            for(int l = 0; l < 3; l++)
                r[3 + 3 * e + l] += be * xi[l];
        }
        for(int pass = 0; pass < 2; pass++)
        {
            int beg = pass == 0 ? rowptr[i] : (crp ? crp[i] : 0);
            int end = pass == 0 ? rowptr[i + 1] : (crp ? crp[i + 1] : 0);
            for(int q = beg; q < end; q++)
            {
                int j = pass == 0 ? cols[q] : crc[q];
                if(agg[j] != c)
                    continue;
                const double* A  = (pass == 0 ? vals : crv) + 9 * (size_t)q;
                double        xj[3] = {x[3 * (size_t)j] - cx, x[3 * (size_t)j + 1] - cy,
                                       x[3 * (size_t)j + 2] - cz};
                acc_jaj(U, A, xi, xj);
            }
        }
    }
    const int t = threadIdx.x;
#pragma unroll
    for(int k = 0; k < 78; k++)
        red[k][t] = U[k];
#pragma unroll
    for(int k = 0; k < 12; k++)
        red[78 + k][t] = r[k];
    __builtin_amdgcn_wave_barrier();  // single wave: LDS accesses stay in order
    for(int row = t; row < 90; row += 64)
    {
        double acc = 0.0;
        for(int l = 0; l < 64; l++)
            acc += red[row][l];
        if(acc != 0.0)
            atomicAdd(row < 78 ? Ac + 78 * (size_t)c + row : rc + 12 * (size_t)c + (row - 78), acc);
    }
}

// Cholesky solve of the packed SPD 12x12 (upper 78); false if not SPD
__device__ __forceinline__ bool chol12_solve(const double* Up, const double* rhs, double* q)
{
    double Lm[78];  // lower factor stored by rows in the same packed order (L^T upper)
    for(int j = 0; j < 12; j++)
    {
        double d = Up[u12(j, j)];
        for(int k = 0; k < j; k++)
            d -= Lm[u12(k, j)] * Lm[u12(k, j)];
        if(!(d > 0.0))
            return false;
        d              = sqrt(d);
        Lm[u12(j, j)]  = d;
        for(int i = j + 1; i < 12; i++)
        {
            double s = Up[u12(j, i)];
            for(int k = 0; k < j; k++)
                s -= Lm[u12(k, j)] * Lm[u12(k, i)];
            Lm[u12(j, i)] = s / d;
        }
    }
    double y[12];
    for(int i = 0; i < 12; i++)  // L y = rhs, L[i][k] = Lm[u12(k, i)]
    {
        double s = rhs[i];
        for(int k = 0; k < i; k++)
            s -= Lm[u12(k, i)] * y[k];
        y[i] = s / Lm[u12(i, i)];
    }
    for(int i = 11; i >= 0; i--)  // L^T q = y
    {
        double s = y[i];
        for(int k = i + 1; k < 12; k++)
            s -= Lm[u12(i, k)] * q[k];
        q[i] = s / Lm[u12(i, i)];
    }
    return true;
}

// per aggregate: q_c, the network features and the scale |q_c|
__global__ void k_agg_solve(int nc, const double* __restrict__ Ac, const double* __restrict__ rc,
                            double* __restrict__ qc, float* __restrict__ feat,
                            double* __restrict__ scale)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= nc)
        return;
    double q[12];
    bool   ok = chol12_solve(Ac + 78 * (size_t)c, rc + 12 * (size_t)c, q);
    double qn = 0.0, rn = 0.0;
    for(int k = 0; k < 12; k++)
    {
        if(!ok)
            q[k] = 0.0;
        qn += q[k] * q[k];
        rn += rc[12 * (size_t)c + k] * rc[12 * (size_t)c + k];
    }
    qn = sqrt(qn);
    rn = sqrt(rn);
    ok = ok && qn > 0.0 && rn > 0.0;
    for(int k = 0; k < 12; k++)
    {
        qc[12 * (size_t)c + k]               = ok ? q[k] : 0.0;
        feat[(size_t)kIn * c + k]            = ok ? (float)(q[k] / qn) : 0.f;
        feat[(size_t)kIn * c + 12 + k]       = ok ? (float)(rc[12 * (size_t)c + k] / rn) : 0.f;
    }
    scale[c] = ok ? qn : 0.0;
}

// q_c += s_c y_c
__global__ void k_agg_correct(int nc, const float* __restrict__ y, const double* __restrict__ scale,
                              double* __restrict__ qc)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= 12 * nc)
        return;
    qc[i] += scale[i / 12] * (double)y[i];
}

// x0_i = J_i q_c for aggregated nodes, zero elsewhere
__global__ void k_agg_prolong(int nn, const int* __restrict__ agg, const double* __restrict__ x,
                              const double* __restrict__ cen, const double* __restrict__ qc,
                              double* __restrict__ x0)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if(v >= nn)
        return;
    int c = agg[v];
    if(c < 0)
    {
        x0[3 * (size_t)v] = x0[3 * (size_t)v + 1] = x0[3 * (size_t)v + 2] = 0.0;
        return;
    }
    const double* q = qc + 12 * (size_t)c;
    double        xb[3];
#pragma unroll
    for(int e = 0; e < 3; e++)
        xb[e] = x[3 * (size_t)v + e] - cen[3 * c + e];
#pragma unroll
    for(int k = 0; k < 3; k++)
        x0[3 * (size_t)v + k] = q[k] + q[3 + 3 * k] * xb[0] + q[4 + 3 * k] * xb[1] + q[5 + 3 * k] * xb[2];
}

// per aggregate: least-squares affine fit q* of a node field (solution)
__global__ void __launch_bounds__(64) k_agg_fit(const int* __restrict__ ptr,
                                                const int* __restrict__ nodes,
                                                const double* __restrict__ x,
                                                const double* __restrict__ cen,
                                                const double* __restrict__ f,
                                                double* __restrict__ qfit, int* __restrict__ okf)
{
    int    c = blockIdx.x;
    double U[78], r[12];
#pragma unroll
    for(int k = 0; k < 78; k++)
        U[k] = 0.0;
#pragma unroll
    for(int k = 0; k < 12; k++)
        r[k] = 0.0;
    const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for(int k = ptr[c] + (int)threadIdx.x; k < ptr[c + 1]; k += 64)
    {
        int    i     = nodes[k];
        double xi[3] = {x[3 * (size_t)i] - cen[3 * c], x[3 * (size_t)i + 1] - cen[3 * c + 1],
                        x[3 * (size_t)i + 2] - cen[3 * c + 2]};
        acc_jaj(U, I3, xi, xi);
#pragma unroll
        for(int e = 0; e < 3; e++)
        {
            double fe = f[3 * (size_t)i + e];
            r[e] += fe;
#pragma unroll
            for(int l = 0; l < 3; l++)
                r[3 + 3 * e + l] += fe * xi[l];
        }
    }
#pragma unroll
    for(int k = 0; k < 78; k++)
        U[k] = wave_sum(U[k]);
#pragma unroll
    for(int k = 0; k < 12; k++)
        r[k] = wave_sum(r[k]);
    if(threadIdx.x == 0)
    {
        double q[12];
        bool   ok = chol12_solve(U, r, q);
        for(int k = 0; k < 12; k++)
            qfit[12 * (size_t)c + k] = ok ? q[k] : 0.0;
        okf[c] = ok ? 1 : 0;
    }
}

// one 128-thread workgroup per sample: h1 = relu(W1^T f + b1), h2 = relu(W2^T
// h1 + b2), y = W3^T h2 + b3.  Optional outputs: hidden layers (training)
__global__ void __launch_bounds__(kHid) k_mlp_forward(int n, MlpW w, const float* __restrict__ X,
                                                      float* __restrict__ Y,
                                                      float* __restrict__ H1,
                                                      float* __restrict__ H2)
{
    __shared__ float f[kIn], h1[kHid], h2[kHid];
    int s = blockIdx.x, t = threadIdx.x;
    if(s >= n)
        return;
    if(t < kIn)
        f[t] = X[(size_t)kIn * s + t];
    __syncthreads();
    float a = w.b1[t];
#pragma unroll
    for(int k = 0; k < kIn; k++)
        a += f[k] * bf16_to_f(w.W1[k * kHid + t]);
    a     = fmaxf(a, 0.f);
    h1[t] = a;
    if(H1)
        H1[(size_t)kHid * s + t] = a;
    __syncthreads();
    float b = w.b2[t];
    for(int j = 0; j < kHid; j++)
        b += h1[j] * bf16_to_f(w.W2[j * kHid + t]);
    b     = fmaxf(b, 0.f);
    h2[t] = b;
    if(H2)
        H2[(size_t)kHid * s + t] = b;
    __syncthreads();
    if(t < kOut)
    {
        float y = w.b3[t];
        for(int j = 0; j < kHid; j++)
            y += h2[j] * bf16_to_f(w.W3[j * kOut + t]);
        Y[(size_t)kOut * s + t] = y;
    }
}

// ------------------------------------------------------------------ training
// dY = 2 (Y - T) / (kOut n); per sample the backward through layers 3 and 2:
// dH2 = (W3 dY) o [h2 > 0], dH1 = (W2 dH2) o [h1 > 0]; loss partials
__global__ void __launch_bounds__(kHid) k_mlp_backward(int n, MlpW w, const float* __restrict__ Y,
                                                       const float* __restrict__ T,
                                                       const float* __restrict__ H1,
                                                       const float* __restrict__ H2,
                                                       float* __restrict__ dY,
                                                       float* __restrict__ dH1,
                                                       float* __restrict__ dH2,
                                                       float* __restrict__ loss)
{
    __shared__ float dy[kOut], d2[kHid];
    int s = blockIdx.x, t = threadIdx.x;
    if(s >= n)
        return;
    if(t < kOut)
    {
        float e = Y[(size_t)kOut * s + t] - T[(size_t)kOut * s + t];
        dy[t]   = 2.f * e / ((float)kOut * (float)n);
        dY[(size_t)kOut * s + t] = dy[t];
        loss[(size_t)kOut * s + t] = e * e / ((float)kOut * (float)n);
    }
    __syncthreads();
    float g = 0.f;
#pragma unroll
    for(int o = 0; o < kOut; o++)
        g += dy[o] * bf16_to_f(w.W3[t * kOut + o]);
    g     = H2[(size_t)kHid * s + t] > 0.f ? g : 0.f;
    d2[t] = g;
    dH2[(size_t)kHid * s + t] = g;
    __syncthreads();
    float h = 0.f;
    for(int j = 0; j < kHid; j++)
        h += d2[j] * bf16_to_f(w.W2[t * kHid + j]);
    dH1[(size_t)kHid * s + t] = H1[(size_t)kHid * s + t] > 0.f ? h : 0.f;
}

// one thread per parameter, summed over the batch in sample order
__global__ void k_mlp_wgrad(int n, const float* __restrict__ X, const float* __restrict__ H1,
                            const float* __restrict__ H2, const float* __restrict__ dY,
                            const float* __restrict__ dH1, const float* __restrict__ dH2,
                            float* __restrict__ grad)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= kParams)
        return;
    const float *A, *G;
    int          lda, ldg, i, j;
    if(p < kOffB1)
    {
        i = p / kHid, j = p % kHid, A = X, lda = kIn, G = dH1, ldg = kHid;
    }
    else if(p < kOffW2)
    {
        i = -1, j = p - kOffB1, A = nullptr, lda = 0, G = dH1, ldg = kHid;
    }
    else if(p < kOffB2)
    {
        int q = p - kOffW2;
        i = q / kHid, j = q % kHid, A = H1, lda = kHid, G = dH2, ldg = kHid;
    }
    else if(p < kOffW3)
    {
        i = -1, j = p - kOffB2, A = nullptr, lda = 0, G = dH2, ldg = kHid;
    }
    else if(p < kOffB3)
    {
        int q = p - kOffW3;
        i = q / kOut, j = q % kOut, A = H2, lda = kHid, G = dY, ldg = kOut;
    }
    else
    {
        i = -1, j = p - kOffB3, A = nullptr, lda = 0, G = dY, ldg = kOut;
    }
    float acc = 0.f;
    for(int s = 0; s < n; s++)
        acc += (A ? A[(size_t)lda * s + i] : 1.f) * G[(size_t)ldg * s + j];
    grad[p] = acc;
}

// Adam on the FP32 master weights, refreshing the BF16 copies of the matrices
__global__ void k_mlp_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, float lr, float b1, float b2, float eps,
                           float c1, float c2, uint16_t* __restrict__ wb)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if(p >= kParams)
        return;
    float gp = g[p];
    float mp = b1 * m[p] + (1.f - b1) * gp;
    float vp = b2 * v[p] + (1.f - b2) * gp * gp;
    m[p]     = mp;
    v[p]     = vp;
    float wp = w[p] - lr * (mp / c1) / (sqrtf(vp / c2) + eps);
    w[p]     = wp;
    wb[p]    = f_to_bf16(wp);
}

__global__ void k_to_bf16(int n, const float* __restrict__ w, uint16_t* __restrict__ wb)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if(p < n)
        wb[p] = f_to_bf16(w[p]);
}

__global__ void k_sum_f(int n, const float* __restrict__ a, double* out)
{
    double s = 0.0;
    for(int i = threadIdx.x; i < n; i += blockDim.x)
        s += (double)a[i];
    s = wave_sum(s);
    __shared__ double ws[4];
    if(__lane_id() == 0)
        ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if(threadIdx.x == 0)
        *out = ws[0] + ws[1] + ws[2] + ws[3];
}

// ---------------------------------------------------------------- host side
MlpW Mlp::view() const
{
    MlpW v;
    v.W1 = wb_.data() + kOffW1;
    v.W2 = wb_.data() + kOffW2;
    v.W3 = wb_.data() + kOffW3;
    v.b1 = w_.data() + kOffB1;
    v.b2 = w_.data() + kOffB2;
    v.b3 = w_.data() + kOffB3;
    return v;
}

void Mlp::init(uint64_t seed, hipStream_t s)
{
    // He-uniform hidden layers, zero output layer: the untrained model is the
    // Galerkin-Jacobi guess
    std::vector<float> h(kParams, 0.f);
    uint64_t           st = seed * 0x9E3779B97F4A7C15ull + 1;
    auto               rnd = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return (double)(st >> 11) * (1.0 / 9007199254740992.0);
    };
    double a1 = std::sqrt(6.0 / kIn), a2 = std::sqrt(6.0 / kHid);
    for(int p = kOffW1; p < kOffB1; p++)
        h[p] = (float)((2 * rnd() - 1) * a1);
    for(int p = kOffW2; p < kOffB2; p++)
        h[p] = (float)((2 * rnd() - 1) * a2);
    set(h.data(), s);
}

void Mlp::set(const float* h, hipStream_t s)
{
    w_.upload(h, kParams, s);
    wb_.resize(kParams);
    m_.resize(kParams);
    v_.resize(kParams);
    m_.zero(s);
    v_.zero(s);
    step_ = 0;
    hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(kParams, 256)), dim3(256), 0, s, kParams,
                       w_.data(), wb_.data());
    SG_LAUNCH_CHECK();
    ready_ = true;
}

void Mlp::get(float* h, hipStream_t s) const
{
    w_.download(h, kParams, s);
    SG_HIP(hipStreamSynchronize(s));
}

void Mlp::ensure_batch(int n)
{
    if(n <= cap_)
        return;
    cap_ = n;
    X_.resize((size_t)kIn * n);
    T_.resize((size_t)kOut * n);
    Y_.resize((size_t)kOut * n);
    dY_.resize((size_t)kOut * n);
    loss_.resize((size_t)kOut * n);
    H1_.resize((size_t)kHid * n);
    H2_.resize((size_t)kHid * n);
    dH1_.resize((size_t)kHid * n);
    dH2_.resize((size_t)kHid * n);
    g_.resize(kParams);
    lsum_.resize(1);
}

double Mlp::loss_grad(const float* X, const float* T, int n, float* grad_host, hipStream_t s)
{
    SG_CHECK(ready_, SGIPC_EINVAL, "network weights not set");
    SG_CHECK(n > 0, SGIPC_EINVAL, "empty batch");
    ensure_batch(n);
    X_.upload(X, (size_t)kIn * n, s);
    T_.upload(T, (size_t)kOut * n, s);
    double L = pass(n, s);
    if(grad_host)
    {
        g_.download(grad_host, kParams, s);
        SG_HIP(hipStreamSynchronize(s));
    }
    return L;
}

// forward + backward on the resident batch; gradient in g_, returns the loss
double Mlp::pass(int n, hipStream_t s)
{
    MlpW w = view();
    hipLaunchKernelGGL(k_mlp_forward, dim3(n), dim3(kHid), 0, s, n, w, X_.data(), Y_.data(),
                       H1_.data(), H2_.data());
    hipLaunchKernelGGL(k_mlp_backward, dim3(n), dim3(kHid), 0, s, n, w, Y_.data(), T_.data(),
                       H1_.data(), H2_.data(), dY_.data(), dH1_.data(), dH2_.data(),
                       loss_.data());
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(grid_for(kParams, 128)), dim3(128), 0, s, n, X_.data(),
                       H1_.data(), H2_.data(), dY_.data(), dH1_.data(), dH2_.data(), g_.data());
    hipLaunchKernelGGL(k_sum_f, dim3(1), dim3(256), 0, s, kOut * n, loss_.data(), lsum_.data());
    SG_LAUNCH_CHECK();
    double L;
    SG_HIP(hipMemcpyAsync(&L, lsum_.data(), sizeof(double), hipMemcpyDeviceToHost, s));
    SG_HIP(hipStreamSynchronize(s));
    return L;
}

double Mlp::train(const float* X, const float* T, int n, int epochs, double lr, hipStream_t s)
{
    SG_CHECK(n > 0 && epochs > 0 && lr > 0.0, SGIPC_EINVAL, "bad training arguments");
    if(!ready_)
        init(1, s);
    ensure_batch(n);
    X_.upload(X, (size_t)kIn * n, s);
    T_.upload(T, (size_t)kOut * n, s);
    double L = 0.0;
    const float b1 = 0.9f, b2 = 0.999f;
    for(int e = 0; e < epochs; e++)
    {
        L = pass(n, s);
        step_++;
        float c1 = 1.f - std::pow(b1, (float)step_), c2 = 1.f - std::pow(b2, (float)step_);
        hipLaunchKernelGGL(k_mlp_adam, dim3(grid_for(kParams, 256)), dim3(256), 0, s, w_.data(),
                           m_.data(), v_.data(), g_.data(), (float)lr, b1, b2, 1e-8f, c1, c2,
                           wb_.data());
        SG_LAUNCH_CHECK();
    }
    return pass(n, s);  // loss of the trained weights
}

void Mlp::infer(const float* feat_dev, int n, float* y_dev, hipStream_t s) const
{
    hipLaunchKernelGGL(k_mlp_forward, dim3(n), dim3(kHid), 0, s, n, view(), feat_dev, y_dev,
                       (float*)nullptr, (float*)nullptr);
    SG_LAUNCH_CHECK();
}

void Coarse::ensure(int nc, int nn)
{
    if(nn > ncap_)
    {
        ncap_ = nn;
        agg_.resize(nn);
        nodes_.resize(nn);
    }
    if(nc <= cap_)
        return;
    cap_ = nc;
    cnt_.resize(nc);
    cur_.resize(nc);
    ptr_.resize(nc + 1);
    cen_.resize(3 * (size_t)nc);
    Ac_.resize(78 * (size_t)nc);
    rc_.resize(12 * (size_t)nc);
    qc_.resize(12 * (size_t)nc);
    qfit_.resize(12 * (size_t)nc);
    okf_.resize(nc);
    scale_.resize(nc);
    feat_.resize((size_t)kIn * nc);
    y_.resize((size_t)kOut * nc);
}

bool Coarse::setup(const Mas& mas, hipStream_t s)
{
    L_ = mas.levels();
    if(L_ < 2 || !mas.ready())
        return false;
    const auto& lv = mas.level_sizes();
    // the coarsest level whose aggregates hold at most kMaxAggregateNodes
    // nodes on average (a connected body is a single node at the coarsest
    // level, which would leave one affine block for the whole body)
    lev_ = L_ - 1;
    while(lev_ > 1 && (long long)lv[2 * lev_] * kMaxAggregateNodes < (long long)lv[0])
        lev_--;
    nc_  = lv[2 * lev_];
    off_ = lv[2 * lev_ + 1];
    chain_         = mas.chain_dev();
    (void)s;
    return nc_ > 0;
}

void Coarse::build(int nn, int nfem, const int* btype, const double* x, const Bsr& A,
                   const Coo* coo, const double* b, hipStream_t s)
{
    ensure(nc_, nn);
    SG_HIP(hipMemsetAsync(cnt_.data(), 0, (size_t)nc_ * sizeof(unsigned), s));
    SG_HIP(hipMemsetAsync(cur_.data(), 0, (size_t)nc_ * sizeof(unsigned), s));
    hipLaunchKernelGGL(k_agg_of, dim3(grid_for(nn, 256)), dim3(256), 0, s, nn, nfem, btype, L_,
                       lev_, chain_, off_, nc_, agg_.data(), cnt_.data());
    std::vector<unsigned> h(nc_);
    cnt_.download(h.data(), nc_, s);
    SG_HIP(hipStreamSynchronize(s));
    std::vector<int> p(nc_ + 1, 0);
    for(int c = 0; c < nc_; c++)
        p[c + 1] = p[c] + (int)h[c];
    ptr_.upload(p, s);
    hipLaunchKernelGGL(k_agg_fill, dim3(grid_for(nn, 256)), dim3(256), 0, s, nn, agg_.data(),
                       ptr_.data(), cur_.data(), nodes_.data());
    hipLaunchKernelGGL(k_agg_centroid, dim3(nc_), dim3(64), 0, s, ptr_.data(), nodes_.data(), x,
                       cen_.data());
    bool rows = coo && coo->rows_valid;
    // enough workgroups per aggregate for about one node per lane
    unsigned hmax = 1;
    for(int c = 0; c < nc_; c++)
        hmax = std::max(hmax, h[c]);
    int G = (int)std::min<unsigned>(64u, (hmax + 63u) / 64u);
    SG_HIP(hipMemsetAsync(Ac_.data(), 0, sizeof(double) * 78 * (size_t)nc_, s));
    SG_HIP(hipMemsetAsync(rc_.data(), 0, sizeof(double) * 12 * (size_t)nc_, s));
    hipLaunchKernelGGL(k_agg_galerkin, dim3(nc_ * G), dim3(64), 0, s, G, ptr_.data(),
                       nodes_.data(), agg_.data(), x, cen_.data(), A.rowptr.data(), A.cols.data(),
                       A.vals.data(), rows ? coo->rp.data() : nullptr,
                       rows ? coo->rc.data() : nullptr, rows ? coo->rv.data() : nullptr, b,
                       Ac_.data(), rc_.data());
    hipLaunchKernelGGL(k_agg_solve, dim3(grid_for(nc_, 64)), dim3(64), 0, s, nc_, Ac_.data(),
                       rc_.data(), qc_.data(), feat_.data(), scale_.data());
    SG_LAUNCH_CHECK();
    nn_ = nn;
    x_  = x;
}

void Coarse::guess(const Mlp* net, double* x0, hipStream_t s)
{
    if(net && net->ready())
    {
        net->infer(feat_.data(), nc_, y_.data(), s);
        hipLaunchKernelGGL(k_agg_correct, dim3(grid_for(12 * nc_, 256)), dim3(256), 0, s, nc_,
                           y_.data(), scale_.data(), qc_.data());
    }
    hipLaunchKernelGGL(k_agg_prolong, dim3(grid_for(nn_, 256)), dim3(256), 0, s, nn_, agg_.data(),
                       x_, cen_.data(), qc_.data(), x0);
    SG_LAUNCH_CHECK();
}

int Coarse::samples(const double* sol, float* feat, float* target, int cap, hipStream_t s)
{
    hipLaunchKernelGGL(k_agg_fit, dim3(nc_), dim3(64), 0, s, ptr_.data(), nodes_.data(), x_,
                       cen_.data(), sol, qfit_.data(), okf_.data());
    SG_LAUNCH_CHECK();
    std::vector<float>  f((size_t)kIn * nc_);
    std::vector<double> qc(12 * (size_t)nc_), qf(12 * (size_t)nc_), sc(nc_);
    std::vector<int>    ok(nc_);
    feat_.download(f.data(), f.size(), s);
    qc_.download(qc.data(), qc.size(), s);
    qfit_.download(qf.data(), qf.size(), s);
    scale_.download(sc.data(), sc.size(), s);
    okf_.download(ok.data(), ok.size(), s);
    SG_HIP(hipStreamSynchronize(s));
    int m = 0;
    for(int c = 0; c < nc_ && m < cap; c++)
    {
        if(sc[c] <= 0.0 || !ok[c])
            continue;
        for(int k = 0; k < kIn; k++)
            feat[(size_t)kIn * m + k] = f[(size_t)kIn * c + k];
        for(int k = 0; k < 12; k++)
            target[(size_t)kOut * m + k] = (float)((qf[12 * c + k] - qc[12 * c + k]) / sc[c]);
        m++;
    }
    return m;
}

}  // namespace sgipc
