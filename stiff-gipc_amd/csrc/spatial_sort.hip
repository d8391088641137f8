// spatial_sort.hip -- spatial re-sorting keys of vertices and tets (the
// reference's re-ordering utilities: _calcVertMChash GIPC.cu:412-461,
// _calcTetMChash :279-341, hash_code :151-198; the remap helpers of
// :343-410 and :9578-9740 are applied by the host to the scene arrays).
//
//   vertices: 30-bit Morton code of the position in the vertices' bounding
//             box (hash_code type -1: the vertex kernel's type stays -1)
//   tets:     the centroid 0.25 ((a + b) + (c + d)) quantised to 10 bits per
//             axis and packed in the axis order the box's extents select
//             (hash_code types 0 .. 5)
// key = code << 32 | index, sorted ascending (this library's radix sort):
// equal codes keep index order.  Built without FMA contraction, so codes
// are bit-identical to a sequential restatement.  The box is the vertices'
// axis-aligned bounding box.  Where the reference leaves a 0/0 coordinate
// of a flat axis to an undefined float-to-integer cast, it is cell 0 here.
#include "sim_kernels.h"

namespace sgipc
{

namespace
{
__device__ __forceinline__ uint32_t dilate3(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// the coordinate scaled to the 1024-cell grid and clamped to [0, 1023]; a
// NaN (an axis of zero extent divides 0 by 0) becomes cell 0
__device__ __forceinline__ double clamp01k(double v)
{
    const double res = 1024.0;
    v                = v * res;
    double a         = (v > 0.0) ? v : 0.0;
    return (res - 1.0 < a) ? res - 1.0 : a;
}

// hash_code(type, x, y, z) on coordinates already divided by the box size
__device__ __forceinline__ uint32_t hash_code(int type, double x, double y, double z)
{
    const uint32_t X = (uint32_t)clamp01k(x), Y = (uint32_t)clamp01k(y), Z = (uint32_t)clamp01k(z);
    switch(type)
    {
        case -1: return (dilate3(X) << 2) + (dilate3(Y) << 1) + dilate3(Z);
        case 0: return (Z * 1024 + Y) * 1024 + X;
        case 1: return (Y * 1024 + Z) * 1024 + X;
        case 2: return (X * 1024 + Z) * 1024 + Y;
        case 3: return (Z * 1024 + X) * 1024 + Y;
        case 4: return (Y * 1024 + X) * 1024 + Z;
        default: return (X * 1024 + Y) * 1024 + Z;
    }
}

// the axis order of the tet key from the box's extents (_calcTetMChash)
__device__ __forceinline__ int tet_hash_type(double sx, double sy, double sz)
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

// bounding box of the vertices: per-block partial max/min, then one block
__global__ void __launch_bounds__(256) k_vbox_partial(int n, const double* __restrict__ x,
                                                      double* __restrict__ part)
{
    __shared__ double s[6][4];
    int    i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[6];
    for(int c = 0; c < 3; c++)
    {
        v[c]     = (i < n) ? x[3 * (size_t)i + c] : -1e32;
        v[3 + c] = (i < n) ? x[3 * (size_t)i + c] : 1e32;
    }
    for(int c = 0; c < 3; c++)
    {
        v[c]     = wave_max(v[c]);
        v[3 + c] = wave_min(v[3 + c]);
    }
    if(__lane_id() == 0)
        for(int c = 0; c < 6; c++)
            s[c][threadIdx.x >> 6] = v[c];
    __syncthreads();
    if(threadIdx.x < 6)
    {
        int    c = threadIdx.x;
        double r = s[c][0];
        for(int w = 1; w < 4; w++)
            r = (c < 3) ? fmax(r, s[c][w]) : fmin(r, s[c][w]);
        part[6 * (size_t)blockIdx.x + c] = r;
    }
}

__global__ void __launch_bounds__(256) k_vbox_final(int nb, const double* __restrict__ part,
                                                    double* __restrict__ box)
{
    __shared__ double s[6][4];
    double v[6] = {-1e32, -1e32, -1e32, 1e32, 1e32, 1e32};
    for(int i = threadIdx.x; i < nb; i += blockDim.x)
        for(int c = 0; c < 6; c++)
        {
            double p = part[6 * (size_t)i + c];
            v[c]     = (c < 3) ? fmax(v[c], p) : fmin(v[c], p);
        }
    for(int c = 0; c < 3; c++)
    {
        v[c]     = wave_max(v[c]);
        v[3 + c] = wave_min(v[3 + c]);
    }
    if(__lane_id() == 0)
        for(int c = 0; c < 6; c++)
            s[c][threadIdx.x >> 6] = v[c];
    __syncthreads();
    if(threadIdx.x < 6)
    {
        int    c = threadIdx.x;
        double r = s[c][0];
        for(int w = 1; w < 4; w++)
            r = (c < 3) ? fmax(r, s[c][w]) : fmin(r, s[c][w]);
        box[c] = r;
    }
}

__global__ void k_vert_keys(int n, const double* __restrict__ x, const double* __restrict__ box,
                            uint64_t* __restrict__ keys)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n)
        return;
    double sx = box[0] - box[3], sy = box[1] - box[4], sz = box[2] - box[5];
    double ox = x[3 * (size_t)i] - box[3], oy = x[3 * (size_t)i + 1] - box[4],
           oz = x[3 * (size_t)i + 2] - box[5];
    uint64_t code = hash_code(-1, ox / sx, oy / sy, oz / sz);
    keys[i]       = (code << 32) | (uint64_t)(uint32_t)i;
}

__global__ void k_tet_keys(int nt, const uint32_t* __restrict__ tets, const double* __restrict__ x,
                           const double* __restrict__ box, uint64_t* __restrict__ keys)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if(e >= nt)
        return;
    double sx = box[0] - box[3], sy = box[1] - box[4], sz = box[2] - box[5];
    const uint32_t* t = tets + 4 * (size_t)e;
    double          c[3];
    for(int k = 0; k < 3; k++)
        c[k] = ((x[3 * (size_t)t[0] + k] + x[3 * (size_t)t[1] + k])
                + (x[3 * (size_t)t[2] + k] + x[3 * (size_t)t[3] + k]))
               * 0.25;
    uint64_t code = hash_code(tet_hash_type(sx, sy, sz), (c[0] - box[3]) / sx,
                              (c[1] - box[4]) / sy, (c[2] - box[5]) / sz);
    keys[e]       = (code << 32) | (uint64_t)(uint32_t)e;
}
}  // namespace

void spatial_sort_keys(int kind, const double* x, int nv, const uint32_t* tets, int nt,
                       uint64_t* keys_out, uint32_t* order_out, hipStream_t s)
{
    SG_CHECK(nv > 0 && x, -1, "no vertices");
    SG_CHECK(kind == 0 || (kind == 1 && nt > 0 && tets), -1, "no tets to sort");
    const int      n = kind == 0 ? nv : nt;
    DBuf<double>   dx, part, box;
    DBuf<uint32_t> dt;
    DBuf<uint64_t> k0, k1;
    DBuf<unsigned> hist;
    dx.upload(x, 3 * (size_t)nv, s);
    const unsigned nb = grid_for(nv, 256);
    part.resize(6 * (size_t)nb);
    box.resize(6);
    hipLaunchKernelGGL(k_vbox_partial, dim3(nb), dim3(256), 0, s, nv, dx.data(), part.data());
    hipLaunchKernelGGL(k_vbox_final, dim3(1), dim3(256), 0, s, (int)nb, part.data(), box.data());
    k0.resize(n);
    k1.resize(n);
    if(kind == 0)
        hipLaunchKernelGGL(k_vert_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, n, dx.data(),
                           box.data(), k0.data());
    else
    {
        for(int e = 0; e < 4 * nt; e++)
            SG_CHECK(tets[e] < (uint32_t)nv, -1, "tet vertex index out of range");
        dt.upload(tets, 4 * (size_t)nt, s);
        hipLaunchKernelGGL(k_tet_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, n, dt.data(),
                           dx.data(), box.data(), k0.data());
    }
    SG_LAUNCH_CHECK();
    hist.resize(radix_sort_hist_size(n));
    radix_sort_keys(k0.data(), k1.data(), hist.data(), n, 32, 62, s);
    std::vector<uint64_t> h = k0.to_host(s);
    if(keys_out)
        std::copy(h.begin(), h.end(), keys_out);
    if(order_out)
        for(int i = 0; i < n; i++)
            order_out[i] = (uint32_t)(h[i] & 0xffffffffu);
}

}  // namespace sgipc
