// noalloc.hip -- sgipc_ls_pcg_noalloc: the PCG solve of sgipc_ls_pcg for
// callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc).
//
// Every device word the solve needs beyond the linear system itself (the
// device copies of b and x, the PCG work vectors, the reduction sets and
// control words, the block-Jacobi inverse) lives in `scratch`, one
// file-scope device array of a fixed size, so a call allocates nothing: no
// hipMalloc, no hipHostMalloc, no event or stream creation.  The kernels are
// the ones sgipc_ls_pcg launches; they receive addresses inside the array.
//
// The host code never passes the array's name to a kernel or a copy: on the
// first call on a device (with that device current) hipGetSymbolAddress gives
// that device's address of its copy of the array, kept per device.
//
// Every call on one GPU shares that GPU's array and leaves the last call's
// contents in it; each call clears what must start at zero (reduction sets,
// control words, x).  Calls that overlap on one GPU (two host threads, or a
// call while another is in flight) race on the array.
#include <chrono>
#include <mutex>

#include "capi.h"

using namespace sgipc;

namespace
{
constexpr int    kMaxDevices = 64;
constexpr size_t kMaxDofs    = 3 * (size_t)SGIPC_NOALLOC_MAX_ROWS;
// b, x and the solver workspace of the largest system (12 reduction sets,
// vectors, block inverses), with margin for the 256-byte rounding of each
// region
constexpr size_t kScratchDoubles = 14 * kMaxDofs + 16 * (size_t)kSlotSet + 4096;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double*    g_scratch[kMaxDevices] = {};

// this device's address of `scratch`, looked up once per device
double* scratch_of_current_device()
{
    int dev = 0;
    SG_HIP(hipGetDevice(&dev));
    SG_CHECK(dev >= 0 && dev < kMaxDevices, SGIPC_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if(!g_scratch[dev])
    {
        void* p = nullptr;
        SG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SG_CHECK(p != nullptr, SGIPC_EHIP, "no device address for the solver scratch array");
        g_scratch[dev] = static_cast<double*>(p);
    }
    return g_scratch[dev];
}

size_t round256(size_t doubles) { return (doubles + 31) / 32 * 32; }
}  // namespace

extern "C" {

int sgipc_ls_pcg_noalloc(sgipc_ls_t ls, const double* b, double* x, const sgipc_pcg_cfg* cfg,
                         sgipc_pcg_stats* stats)
{
    return guard([&] {
        SG_CHECK(ls && b && x && cfg, SGIPC_EINVAL, "null argument");
        SG_CHECK(cfg->solver_precision == SGIPC_SOLVE_FP64, SGIPC_EINVAL,
                 "sgipc_ls_pcg_noalloc: FP64 solves only (the mixed-precision solve keeps an "
                 "FP32 copy of the matrix, which is allocated)");
        SG_CHECK(ls->A.nb <= SGIPC_NOALLOC_MAX_ROWS, SGIPC_ECAPACITY,
                 "sgipc_ls_pcg_noalloc: " + std::to_string(ls->A.nb)
                     + " block rows exceed the maximum of "
                     + std::to_string(SGIPC_NOALLOC_MAX_ROWS));
        int dev = 0;
        SG_HIP(hipGetDevice(&dev));
        SG_CHECK(dev == ls->dev, SGIPC_EINVAL,
                 "sgipc_ls_pcg_noalloc: the current device is not the system's device");
        const int n    = 3 * ls->A.nb;
        double*   base = scratch_of_current_device();
        double*   bd   = base;
        double*   xd   = bd + round256(n);
        double*   ws   = xd + round256(n);
        SG_CHECK((size_t)(ws - base) + Pcg::workspace_doubles(n) <= kScratchDoubles, SGIPC_ECAPACITY,
                 "sgipc_ls_pcg_noalloc: workspace exceeds the scratch array");
        ls->pcg_na.bind_workspace(ws, n);

        PcgConfig pc;
        pc.rel_tol     = cfg->rel_tol;
        pc.abs_tol     = cfg->abs_tol;
        pc.max_iter    = cfg->max_iter;
        pc.use_rz      = cfg->use_preconditioned_norm != 0;
        pc.check_every = cfg->check_every;
        pc.mixed       = false;
        auto t0        = std::chrono::steady_clock::now();
        SG_HIP(hipMemcpyAsync(bd, b, n * sizeof(double), hipMemcpyHostToDevice, ls->s));
        PcgStats st = ls->pcg_na.solve(SysView{&ls->A, nullptr},
                                       ls->precond == SGIPC_PRECOND_MAS ? Precond::Mas
                                                                        : Precond::Diag,
                                       &ls->mas, bd, xd, pc, ls->s);
        SG_HIP(hipMemcpyAsync(x, xd, n * sizeof(double), hipMemcpyDeviceToHost, ls->s));
        SG_HIP(hipStreamSynchronize(ls->s));
        auto t1 = std::chrono::steady_clock::now();
        if(stats)
        {
            stats->iterations       = st.iterations;
            stats->outer_iterations = st.outer_iterations;
            stats->converged        = st.converged ? 1 : 0;
            stats->rel_residual     = st.rel_residual;
            stats->rz0              = st.rz0;
            stats->rz               = st.rz;
            stats->ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        }
    });
}

}  // extern "C"
