// capi.h -- pieces shared by the translation units of the extern "C"
// boundary: the status/message guard and the stand-alone linear system.
#pragma once

#include <string>

#include "linsys.h"
#include "sgipc.h"

namespace sgipc
{
// message of the last failure on this thread (sgipc_last_error)
extern thread_local std::string g_err;

// run f, converting exceptions into a status code and g_err
template <typename F>
int guard(F&& f)
{
    try
    {
        f();
        g_err.clear();
        return SGIPC_OK;
    }
    catch(const Error& e)
    {
        g_err = e.what();
        return e.code;
    }
    catch(const std::exception& e)
    {
        g_err = e.what();
        return SGIPC_EHIP;
    }
    catch(...)
    {
        g_err = "unknown error";
        return SGIPC_EHIP;
    }
}
}  // namespace sgipc

struct sgipc_linear_system
{
    hipStream_t          s       = nullptr;  // stream of the system's work
    hipStream_t          own     = nullptr;  // the stream it created
    int                  dev     = 0;  // device current at creation
    sgipc::Bsr           A;
    sgipc::Mas           mas;
    sgipc::Pcg           pcg;
    sgipc::Pcg           pcg_na;       // solver of sgipc_ls_pcg_noalloc (no memory of its own)
    int                  precond = 1;
    sgipc::DBuf<double>  b, x, y;
    ~sgipc_linear_system()
    {
        if(own)
            (void)hipStreamDestroy(own);
    }
};
