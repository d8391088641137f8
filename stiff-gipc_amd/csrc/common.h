// common.h -- HIP runtime helpers shared by every translation unit of
// libsgipc.so: error propagation, device buffers, launch shapes and wave64
// reductions.  Written for gfx950 (CDNA4, 64-lane wavefronts).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace sgipc
{

// ---------------------------------------------------------------- errors
// The reference aborts on every CUDA error (cuda_tools.h:172-206); this
// library throws, and the C ABI converts the exception into a status code
// plus a thread-local message.
struct Error : std::runtime_error
{
    int code;
    Error(int c, const std::string& m)
        : std::runtime_error(m)
        , code(c)
    {
    }
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);

#define SG_HIP(expr)                                                              \
    do                                                                            \
    {                                                                             \
        hipError_t _e = (expr);                                                   \
        if(_e != hipSuccess)                                                      \
            ::sgipc::throw_hip(_e, #expr, __FILE__, __LINE__);                    \
    } while(0)

#define SG_CHECK(cond, code, msg)                                                 \
    do                                                                            \
    {                                                                             \
        if(!(cond))                                                               \
            throw ::sgipc::Error((code), std::string(msg));                       \
    } while(0)

#define SG_LAUNCH_CHECK() SG_HIP(hipGetLastError())

// ---------------------------------------------------------------- buffers
template <typename T>
class DBuf
{
  public:
    DBuf() = default;
    explicit DBuf(size_t n) { resize(n); }
    ~DBuf() { release(); }
    DBuf(const DBuf&)            = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept
        : p_(o.p_)
        , n_(o.n_)
        , cap_(o.cap_)
        , owned_(o.owned_)
    {
        o.p_     = nullptr;
        o.n_     = 0;
        o.cap_   = 0;
        o.owned_ = true;
    }
    DBuf& operator=(DBuf&& o) noexcept
    {
        if(this != &o)
        {
            release();
            p_       = o.p_;
            n_       = o.n_;
            cap_     = o.cap_;
            owned_   = o.owned_;
            o.p_     = nullptr;
            o.n_     = 0;
            o.cap_   = 0;
            o.owned_ = true;
        }
        return *this;
    }
    // grow-only reallocation; contents are not preserved.  A borrowed buffer
    // never reallocates: a size beyond its capacity is an error.
    void resize(size_t n)
    {
        if(n > cap_)
        {
            SG_CHECK(owned_, -3, "a borrowed device buffer is too small");
            release();
            size_t bytes = (n ? n : 1) * sizeof(T);
            SG_HIP(hipMalloc(&p_, bytes));
            cap_ = n ? n : 1;
        }
        n_ = n;
    }
    // use n elements of memory owned elsewhere (never freed here)
    void borrow(T* p, size_t n)
    {
        release();
        p_     = p;
        n_     = n;
        cap_   = n;
        owned_ = false;
    }
    void release()
    {
        if(p_ && owned_)
            (void)hipFree(p_);
        p_     = nullptr;
        n_     = 0;
        cap_   = 0;
        owned_ = true;
    }
    void zero(hipStream_t s = 0)
    {
        if(n_)
            SG_HIP(hipMemsetAsync(p_, 0, n_ * sizeof(T), s));
    }
    void upload(const T* h, size_t n, hipStream_t s = 0)
    {
        resize(n);
        if(n)
            SG_HIP(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& h, hipStream_t s = 0) { upload(h.data(), h.size(), s); }
    void download(T* h, size_t n, hipStream_t s = 0) const
    {
        if(n)
            SG_HIP(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    std::vector<T> to_host(hipStream_t s = 0) const
    {
        std::vector<T> h(n_);
        download(h.data(), n_, s);
        SG_HIP(hipStreamSynchronize(s));
        return h;
    }
    T*       data() { return p_; }
    const T* data() const { return p_; }
    size_t   size() const { return n_; }

  private:
    T*     p_     = nullptr;
    size_t n_     = 0;
    size_t cap_   = 0;
    bool   owned_ = true;
};

// ---------------------------------------------------------------- profiler
// Event pairs recorded around selected launches on the launching stream;
// resolved lazily (after the stream has been synchronised).
class Profiler
{
  public:
    bool enabled = false;
    struct Entry
    {
        int    launches = 0;
        double total_ms = 0.0;
        long   calls    = 0;
        bool   open     = false;
        std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    };
    ~Profiler()
    {
        for(auto& kv : e_)
            for(auto& p : kv.second.pending)
                destroy(p);
        for(auto& p : pool_)
            destroy(p);
    }
    // every sample-th launch of each name is timed (an event pair costs a few
    // microseconds of stream time, so sparse sampling keeps timed runs honest)
    int sample = 1;
    void begin(const char* name, hipStream_t s)
    {
        if(!enabled)
            return;
        auto& ent = e_[name];
        ent.open  = (ent.calls++ % (sample > 0 ? sample : 1)) == 0;
        if(!ent.open)
            return;
        auto p = get();
        ent.pending.push_back(p);
        SG_HIP(hipEventRecord(p.first, s));
    }
    // events for hipExtLaunchKernel, which records them with the dispatch
    // itself: they bracket the kernel's own execution (what a kernel-trace
    // profiler reports), without a separate marker packet before and after.
    // Nulls when profiling is off or this launch is not sampled.
    std::pair<hipEvent_t, hipEvent_t> kernel_events(const char* name)
    {
        if(!enabled)
            return {nullptr, nullptr};
        auto& ent = e_[name];
        if((ent.calls++ % (sample > 0 ? sample : 1)) != 0)
            return {nullptr, nullptr};
        auto p = get();
        ent.pending.push_back(p);
        return p;
    }
    void end(const char* name, hipStream_t s)
    {
        if(!enabled)
            return;
        auto& ent = e_[name];
        if(!ent.open)
            return;
        ent.open = false;
        SG_HIP(hipEventRecord(ent.pending.back().second, s));
    }
    // call after the stream has been synchronised
    void resolve()
    {
        for(auto& kv : e_)
        {
            for(auto& p : kv.second.pending)
            {
                float ms = 0;
                SG_HIP(hipEventElapsedTime(&ms, p.first, p.second));
                kv.second.total_ms += ms;
                kv.second.launches++;
                pool_.push_back(p);
            }
            kv.second.pending.clear();
        }
    }
    void clear()
    {
        for(auto& kv : e_)
        {
            for(auto& p : kv.second.pending)
                pool_.push_back(p);
        }
        e_.clear();
    }
    Entry query(const std::string& n)
    {
        auto it = e_.find(n);
        Entry r;
        if(it != e_.end())
        {
            r.launches = it->second.launches;
            r.total_ms = it->second.total_ms;
        }
        return r;
    }

  private:
    std::pair<hipEvent_t, hipEvent_t> get()
    {
        if(!pool_.empty())
        {
            auto p = pool_.back();
            pool_.pop_back();
            return p;
        }
        std::pair<hipEvent_t, hipEvent_t> p;
        SG_HIP(hipEventCreate(&p.first));
        SG_HIP(hipEventCreate(&p.second));
        return p;
    }
    static void destroy(std::pair<hipEvent_t, hipEvent_t>& p)
    {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    std::map<std::string, Entry>                   e_;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool_;
};

// ---------------------------------------------------------------- launch
constexpr int kWave = 64;

inline unsigned grid_for(long long n, int block)
{
    long long g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

// The command processor hands workgroups of a 1-D grid to the 8 XCDs round
// robin (workgroup b to XCD b % 8), and each XCD has its own L2.  This maps
// the hardware workgroup id to a logical one so that each XCD works on one
// contiguous range of logical workgroups: XCD k takes logical ids
// [k q + min(k, r), (k + 1) q + min(k + 1, r)) with q = G / 8, r = G % 8,
// a bijection of [0, G).  Kernels whose neighbouring rows share operands
// (SpMV's x gathers) then read each operand into one L2 instead of eight.
constexpr int kXcds = 8;

__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned G)
{
    unsigned q = G / kXcds, r = G % kXcds;
    unsigned k = b % kXcds, i = b / kXcds;
    return k * q + (k < r ? k : r) + i;
}

// ---------------------------------------------------------------- hashing
// open-addressing tables keyed by 64-bit (row << 32 | col)
constexpr unsigned long long kEmptyKey = ~0ull;

__host__ __device__ __forceinline__ unsigned long long mix64(unsigned long long k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

// ---------------------------------------------------------------- wave ops
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for(int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for(int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for(int o = 32; o > 0; o >>= 1)
    {
        double u = __shfl_xor(v, o, 64);
        v        = (v < u) ? u : v;
    }
    return v;
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for(int o = 32; o > 0; o >>= 1)
    {
        double u = __shfl_xor(v, o, 64);
        v        = (u < v) ? u : v;
    }
    return v;
}

// positive doubles order like their bit patterns: atomic max via u64 max
__device__ __forceinline__ void atomic_max_pos(double* addr, double v)
{
    atomicMax(reinterpret_cast<unsigned long long*>(addr), (unsigned long long)__double_as_longlong(v));
}

// ------------------------------------------------------------ slot sets
// A global FP64 sum over the workgroups of one launch, in a fixed order so
// that reruns give identical bits.  Each workgroup reduces its values (wave
// butterflies, then the waves in index order) and stores the result in its
// own partial; a one-workgroup launch that follows in stream order
// (set_finalize) sums the partials in a fixed order and adds the total to
// slot 0.  A reduction across workgroups inside the launch itself would need
// an agent-scope release fence per workgroup, and each of those writes back
// the L2: measured, it tripled the PCG SpMV's time (170 against 55 us on a
// 349k-row system), where the extra launch costs a few microseconds.
//
// Layout of a set (doubles): kSlots result slots (slot 0 holds the sum, the
// others stay 0: a consumer wave sums all kSlots, one per lane, so a set is
// read as 512 contiguous bytes), then the workgroup partials.  Launches may
// add to the same set one after another (each with its finalize; the total
// accumulates in slot 0); a launch may have at most kSetMaxBlocks workgroups.
constexpr int kSlots        = 64;
constexpr int kSlotStride   = 1;
constexpr int kSetMaxBlocks = 65536;
constexpr int kSetOffBlock  = kSlots;
constexpr int kSlotSet      = kSetOffBlock + kSetMaxBlocks;

// every lane of the calling wave receives the total of the set
__device__ __forceinline__ double slot_set_sum(const double* set)
{
    return wave_sum(set[__lane_id() * kSlotStride]);
}

// this workgroup's sum of v stored as its partial of the set; call from every
// thread of every workgroup of the launch (contains __syncthreads);
// blockDim.x a multiple of 64, <= 1024; gridDim.x <= kSetMaxBlocks
__device__ __forceinline__ void slot_set_add(double* set, double v)
{
    __shared__ double part[16];
    v           = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if(__lane_id() == 0)
        part[w] = v;
    __syncthreads();
    if(threadIdx.x == 0)
    {
        double t  = 0.0;
        int    nw = blockDim.x >> 6;
        for(int i = 0; i < nw; i++)
            t += part[i];
        set[kSetOffBlock + blockIdx.x] = t;
    }
}

// the total of a set whose nparts workgroup partials have not been
// finalized, computed by one wave in a fixed order: lane i loads partials
// i + 64 j, j < 16, all in one round (1024 partials per round), adds them in
// four interleaved chains (j mod 4), then the butterfly; every lane of the
// wave receives it.  A consumer launch can finalize a set this way instead of
// a set_finalize launch before it.
__device__ __forceinline__ double slot_set_partials_sum(const double* set, int nparts)
{
    const double* p    = set + kSetOffBlock;
    const int     lane = (int)__lane_id();
    double        c[4] = {0.0, 0.0, 0.0, 0.0};
    for(int b = 0; b < nparts; b += 1024)
    {
        double v[16];
#pragma unroll
        for(int q = 0; q < 16; q++)
        {
            const int i = b + 64 * q + lane;
            v[q]        = i < nparts ? p[i] : 0.0;
        }
#pragma unroll
        for(int q = 0; q < 16; q++)
            c[q & 3] += v[q];
    }
    return wave_sum((c[0] + c[1]) + (c[2] + c[3]));
}

// set_finalize's body: one workgroup of 1024 threads sums the nblocks
// partials: thread t adds partials t + 1024 j in four interleaved chains
// (j mod 4), whose loads are all issued before the adds, then the chains,
// the lanes (butterfly) and the waves, each in a fixed order
constexpr int kFinalizeThreads = 1024;
__device__ __forceinline__ void slot_set_finalize(double* set, int nblocks)
{
    __shared__ double part[kFinalizeThreads / 64];
    const double*     p = set + kSetOffBlock;
    double            c[4] = {0.0, 0.0, 0.0, 0.0};
    for(int b = 0; b < nblocks; b += 4 * kFinalizeThreads)
    {
        double v[4];
#pragma unroll
        for(int q = 0; q < 4; q++)
        {
            const int i = b + q * kFinalizeThreads + (int)threadIdx.x;
            v[q]        = i < nblocks ? p[i] : 0.0;
        }
#pragma unroll
        for(int q = 0; q < 4; q++)
            c[q] += v[q];
    }
    double t = wave_sum((c[0] + c[1]) + (c[2] + c[3]));
    if(__lane_id() == 0)
        part[threadIdx.x >> 6] = t;
    __syncthreads();
    if(threadIdx.x == 0)
    {
        double tot = 0.0;
        for(int w = 0; w < kFinalizeThreads / 64; w++)
            tot += part[w];
        set[0] = set[0] + tot;
    }
}

// clear the result slots of one set (kSlots threads of any workgroup)
__device__ __forceinline__ void slot_set_clear(double* set, int lane)
{
    if(lane < kSlots)
        set[lane * kSlotStride] = 0.0;
}

// wave-aggregated append: one atomic per wave, returns this lane's slot
// (only lanes with pred != 0 receive a meaningful slot)
__device__ __forceinline__ unsigned wave_append(unsigned* counter, bool pred)
{
    unsigned long long m    = __ballot(pred);
    unsigned           lane = __lane_id();
    unsigned           base = 0;
    int                leader = __ffsll((long long)m) - 1;
    if(m && (int)lane == leader)
        base = atomicAdd(counter, (unsigned)__popcll(m));
    base = __shfl(base, leader < 0 ? 0 : leader, 64);
    unsigned long long lt = (lane == 0) ? 0ull : (m & ((~0ull) >> (64 - lane)));
    return base + (unsigned)__popcll(lt);
}

// wave-aggregated append into per-key counters: one atomic per distinct key
// in the wave (lanes with key < 0 take no slot); returns this lane's slot
// under counters[key]
__device__ __forceinline__ unsigned wave_append_keyed(unsigned* counters, int key)
{
    const unsigned           lane = __lane_id();
    const unsigned long long lt   = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    unsigned                 mine = 0;
    unsigned long long       todo = __ballot(key >= 0);
    while(todo)
    {
        int                leader = __ffsll((long long)todo) - 1;
        int                k      = __shfl(key, leader, 64);
        unsigned long long m      = __ballot(key == k) & todo;
        unsigned           base   = 0;
        if((int)lane == leader)
            base = atomicAdd(counters + k, (unsigned)__popcll(m));
        base = __shfl(base, leader, 64);
        if(key == k)
            mine = base + (unsigned)__popcll(m & lt);
        todo &= ~m;
    }
    return mine;
}

}  // namespace sgipc
