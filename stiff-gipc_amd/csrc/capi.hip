// capi.hip -- extern "C" boundary of libsgipc.so (include/sgipc.h).
// Every entry point converts C++ exceptions into a status code and a
// thread-local message.
#include <chrono>
#include <cstring>

#include "../../include/sgipc.h"
#include "capi.h"
#include "context.h"
#include "mm_io.h"

using namespace sgipc;

namespace sgipc
{
thread_local std::string g_err;
}
extern "C" {

const char* sgipc_last_error(void) { return g_err.c_str(); }

int sgipc_device_info(char* buf, int32_t n)
{
    return guard([&] {
        int dev = 0;
        SG_HIP(hipGetDevice(&dev));
        hipDeviceProp_t p;
        SG_HIP(hipGetDeviceProperties(&p, dev));
        snprintf(buf, n, "%s %s CUs=%d", p.gcnArchName, p.name, p.multiProcessorCount);
    });
}

int sgipc_ls_create(int32_t nb, int32_t nnzb, const int32_t* rowptr, const int32_t* cols,
                    const double* vals, const int32_t* part_id, int32_t precond,
                    int32_t mas_precision, sgipc_ls_t* out)
{
    return guard([&] {
        SG_CHECK(out && rowptr && cols && vals && nb > 0 && nnzb > 0, SGIPC_EINVAL,
                 "sgipc_ls_create: null argument or empty matrix");
        SG_CHECK(rowptr[0] == 0 && rowptr[nb] == nnzb, SGIPC_EINVAL,
                 "sgipc_ls_create: rowptr must start at 0 and end at nnzb");
        for(int i = 0; i < nb; i++)
        {
            SG_CHECK(rowptr[i + 1] >= rowptr[i], SGIPC_EINVAL, "rowptr not monotone");
            for(int k = rowptr[i]; k < rowptr[i + 1]; k++)
            {
                SG_CHECK(cols[k] >= 0 && cols[k] < nb, SGIPC_EINVAL, "column out of range");
                SG_CHECK(k == rowptr[i] || cols[k] > cols[k - 1], SGIPC_EINVAL,
                         "columns must be strictly increasing within a row");
            }
        }
        SG_CHECK(precond == SGIPC_PRECOND_DIAG || precond == SGIPC_PRECOND_MAS, SGIPC_EINVAL,
                 "unknown preconditioner");
        auto* ls = new sgipc_linear_system();
        try
        {
            SG_HIP(hipGetDevice(&ls->dev));
            SG_HIP(hipStreamCreateWithFlags(&ls->s, hipStreamNonBlocking));
            ls->own = ls->s;
            std::vector<int> rp(rowptr, rowptr + nb + 1), cl(cols, cols + nnzb);
            ls->A.set_pattern(nb, rp, cl, ls->s);
            ls->A.vals.upload(vals, (size_t)9 * nnzb, ls->s);
            ls->precond = precond;
            if(precond == SGIPC_PRECOND_MAS)
            {
                MasConfig mc;
                mc.fp32 = (mas_precision == SGIPC_MAS_FP32);
                mc.blocked_invert = mas_blocked_invert();
                ls->mas.init(ls->A, part_id, mc, ls->s);
                ls->mas.build(SysView{&ls->A, nullptr}, ls->s);
            }
            else
                ls->pcg.build_diag(ls->A, ls->s);
            SG_HIP(hipStreamSynchronize(ls->s));
        }
        catch(...)
        {
            delete ls;
            throw;
        }
        *out = ls;
    });
}

int sgipc_ls_create_bcoo(int32_t nb, int32_t nnz, const int32_t* rows, const int32_t* cols,
                         const double* vals, const int32_t* part_id, int32_t precond,
                         int32_t mas_precision, sgipc_ls_t* out)
{
    std::vector<int>    rp, cl;
    std::vector<double> vl;
    int rc = guard([&] {
        SG_CHECK(rows && cols && vals && nb > 0 && nnz > 0, SGIPC_EINVAL,
                 "sgipc_ls_create_bcoo: null argument or empty matrix");
        upper_bcoo_to_full_bsr(nb, nnz, rows, cols, vals, rp, cl, vl);
    });
    if(rc != SGIPC_OK)
        return rc;
    return sgipc_ls_create(nb, (int32_t)cl.size(), rp.data(), cl.data(), vl.data(), part_id,
                           precond, mas_precision, out);
}

int sgipc_ls_create_mm(const char* A_path, const int32_t* part_id, int32_t precond,
                       int32_t mas_precision, sgipc_ls_t* out)
{
    int                 nb = 0;
    std::vector<int>    br, bc;
    std::vector<double> bv;
    int rc = guard([&] {
        SG_CHECK(A_path, SGIPC_EINVAL, "null path");
        mm_read_upper_blocks(A_path, &nb, br, bc, bv);
    });
    if(rc != SGIPC_OK)
        return rc;
    return sgipc_ls_create_bcoo(nb, (int32_t)br.size(), br.data(), bc.data(), bv.data(), part_id,
                                precond, mas_precision, out);
}

int sgipc_mm_read_upper_bcoo(const char* A_path, int32_t* nb, int32_t* nnz, int32_t* rows,
                             int32_t* cols, double* vals, int32_t cap)
{
    return guard([&] {
        SG_CHECK(A_path && nb && nnz, SGIPC_EINVAL, "null argument");
        int                 n = 0;
        std::vector<int>    br, bc;
        std::vector<double> bv;
        mm_read_upper_blocks(A_path, &n, br, bc, bv);
        *nb       = n;
        *nnz      = (int32_t)br.size();
        size_t m  = std::min<size_t>(br.size(), cap > 0 ? (size_t)cap : 0);
        if(rows)
            std::copy(br.begin(), br.begin() + m, rows);
        if(cols)
            std::copy(bc.begin(), bc.begin() + m, cols);
        if(vals)
            std::copy(bv.begin(), bv.begin() + 9 * m, vals);
    });
}

int sgipc_ls_block_rows(sgipc_ls_t ls, int32_t* nb)
{
    return guard([&] {
        SG_CHECK(ls && nb, SGIPC_EINVAL, "null argument");
        *nb = ls->A.nb;
    });
}

int sgipc_mm_read_vector(const char* path, double* b, int32_t cap, int32_t* m)
{
    return guard([&] {
        SG_CHECK(path && m, SGIPC_EINVAL, "null argument");
        std::vector<double> v = mm_read_vector(path);
        *m                    = (int32_t)v.size();
        if(b && cap > 0)
            std::copy(v.begin(), v.begin() + std::min<size_t>(v.size(), (size_t)cap), b);
    });
}

int sgipc_spatial_sort_keys(int32_t kind, const double* x, int32_t n_verts,
                            const uint32_t* tets, int32_t n_tets, uint64_t* keys,
                            uint32_t* order)
{
    return guard([&] {
        SG_CHECK(kind == SGIPC_SORT_VERTICES || kind == SGIPC_SORT_TETS, SGIPC_EINVAL,
                 "kind must be SGIPC_SORT_VERTICES or SGIPC_SORT_TETS");
        SG_CHECK(x && n_verts > 0, SGIPC_EINVAL, "no vertices");
        SG_CHECK(kind == SGIPC_SORT_VERTICES || (tets && n_tets > 0), SGIPC_EINVAL, "no tets");
        hipStream_t s;
        SG_HIP(hipStreamCreate(&s));
        try
        {
            spatial_sort_keys(kind, x, n_verts, tets, n_tets, keys, order, s);
        }
        catch(...)
        {
            (void)hipStreamDestroy(s);
            throw;
        }
        SG_HIP(hipStreamDestroy(s));
    });
}

int sgipc_ls_set_stream(sgipc_ls_t ls, void* stream)
{
    return guard([&] {
        SG_CHECK(ls, SGIPC_EINVAL, "null system");
        SG_HIP(hipStreamSynchronize(ls->s));
        ls->s = stream ? static_cast<hipStream_t>(stream) : ls->own;
    });
}

int sgipc_ls_destroy(sgipc_ls_t ls)
{
    return guard([&] { delete ls; });
}

int sgipc_ls_spmv(sgipc_ls_t ls, const double* x, double* y)
{
    return guard([&] {
        SG_CHECK(ls && x && y, SGIPC_EINVAL, "null argument");
        size_t n = 3 * (size_t)ls->A.nb;
        ls->x.upload(x, n, ls->s);
        ls->y.resize(n);
        spmv(SysView{&ls->A, nullptr}, ls->x.data(), ls->y.data(), nullptr, ls->s);
        ls->y.download(y, n, ls->s);
        SG_HIP(hipStreamSynchronize(ls->s));
    });
}

int sgipc_ls_precond_apply(sgipc_ls_t ls, const double* r, double* z)
{
    return guard([&] {
        SG_CHECK(ls && r && z, SGIPC_EINVAL, "null argument");
        size_t n = 3 * (size_t)ls->A.nb;
        ls->x.upload(r, n, ls->s);
        ls->y.resize(n);
        if(ls->precond == SGIPC_PRECOND_MAS)
            ls->mas.apply(ls->x.data(), ls->y.data(), ls->s);
        else
            ls->pcg.apply_diag(ls->A.nb, ls->x.data(), ls->y.data(), ls->s);
        ls->y.download(z, n, ls->s);
        SG_HIP(hipStreamSynchronize(ls->s));
    });
}

int sgipc_ls_pcg(sgipc_ls_t ls, const double* b, double* x, const sgipc_pcg_cfg* cfg,
                 sgipc_pcg_stats* stats)
{
    return guard([&] {
        SG_CHECK(ls && b && x && cfg, SGIPC_EINVAL, "null argument");
        SG_CHECK(cfg->solver_precision == SGIPC_SOLVE_FP64
                     || cfg->solver_precision == SGIPC_SOLVE_MIXED,
                 SGIPC_EINVAL, "unknown solver precision");
        size_t n = 3 * (size_t)ls->A.nb;
        ls->b.upload(b, n, ls->s);
        ls->x.resize(n);
        PcgConfig pc;
        pc.rel_tol     = cfg->rel_tol;
        pc.abs_tol     = cfg->abs_tol;
        pc.max_iter    = cfg->max_iter;
        pc.use_rz      = cfg->use_preconditioned_norm != 0;
        pc.check_every = cfg->check_every;
        pc.mixed       = cfg->solver_precision == SGIPC_SOLVE_MIXED;
        hipEvent_t e0, e1;
        SG_HIP(hipEventCreate(&e0));
        SG_HIP(hipEventCreate(&e1));
        SG_HIP(hipEventRecord(e0, ls->s));
        PcgStats st = ls->pcg.solve(SysView{&ls->A, nullptr},
                                    ls->precond == SGIPC_PRECOND_MAS ? Precond::Mas : Precond::Diag,
                                    &ls->mas, ls->b.data(), ls->x.data(), pc, ls->s);
        SG_HIP(hipEventRecord(e1, ls->s));
        ls->x.download(x, n, ls->s);
        SG_HIP(hipStreamSynchronize(ls->s));
        float ms = 0;
        SG_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if(stats)
        {
            stats->iterations       = st.iterations;
            stats->outer_iterations = st.outer_iterations;
            stats->converged        = st.converged ? 1 : 0;
            stats->rel_residual     = st.rel_residual;
            stats->rz0              = st.rz0;
            stats->rz               = st.rz;
            stats->ms               = ms;
        }
    });
}

int sgipc_ls_mas_info(sgipc_ls_t ls, int32_t* levels, int32_t* level_sizes,
                      int32_t* total_clusters)
{
    return guard([&] {
        SG_CHECK(ls && ls->precond == SGIPC_PRECOND_MAS && ls->mas.ready(), SGIPC_ESTATE,
                 "no MAS hierarchy on this system");
        if(levels)
            *levels = ls->mas.levels();
        if(total_clusters)
            *total_clusters = ls->mas.total_clusters();
        if(level_sizes)
        {
            const auto& v = ls->mas.level_sizes();
            for(size_t i = 0; i < v.size(); i++)
                level_sizes[i] = v[i];
        }
    });
}

int sgipc_ls_mas_chain(sgipc_ls_t ls, int32_t* chain)
{
    return guard([&] {
        SG_CHECK(ls && ls->precond == SGIPC_PRECOND_MAS && ls->mas.ready() && chain,
                 SGIPC_ESTATE, "no MAS hierarchy on this system");
        auto v = ls->mas.chain_host(ls->s);
        std::memcpy(chain, v.data(), v.size() * sizeof(int));
    });
}

int sgipc_ls_mas_bank(sgipc_ls_t ls, int32_t bank, double* mat, double* inv)
{
    return guard([&] {
        SG_CHECK(ls && ls->precond == SGIPC_PRECOND_MAS && ls->mas.ready(), SGIPC_ESTATE,
                 "no MAS hierarchy on this system");
        ls->mas.bank_host(bank, mat, inv, ls->s);
    });
}

// ------------------------------------------------------------------ context
struct sgipc_context
{
    Context* c = nullptr;
    ~sgipc_context() { delete c; }
};

static Context& ctx_of(sgipc_t h)
{
    SG_CHECK(h && h->c, SGIPC_EINVAL, "null context");
    return *h->c;
}

int sgipc_create(const sgipc_scene_desc* desc, sgipc_t* out)
{
    return guard([&] {
        SG_CHECK(desc && out, SGIPC_EINVAL, "null argument");
        auto* h = new sgipc_context();
        try
        {
            h->c = new Context(*desc);
        }
        catch(...)
        {
            delete h;
            throw;
        }
        *out = h;
    });
}

int sgipc_destroy(sgipc_t h)
{
    return guard([&] { delete h; });
}

int sgipc_set_positions(sgipc_t h, const double* x)
{
    return guard([&] {
        SG_CHECK(x, SGIPC_EINVAL, "null positions");
        ctx_of(h).set_positions(x);
    });
}

int sgipc_get_positions(sgipc_t h, double* x)
{
    return guard([&] {
        SG_CHECK(x, SGIPC_EINVAL, "null output");
        auto& c = ctx_of(h);
        c.get(c.x_, x);
    });
}

int sgipc_get_velocities(sgipc_t h, double* v)
{
    return guard([&] {
        SG_CHECK(v, SGIPC_EINVAL, "null output");
        auto& c = ctx_of(h);
        c.get(c.v_, v);
    });
}

int sgipc_set_xtilde(sgipc_t h, const double* xt)
{
    return guard([&] {
        SG_CHECK(xt, SGIPC_EINVAL, "null x_tilde");
        ctx_of(h).set_xtilde(xt);
    });
}

int sgipc_get_xtilde(sgipc_t h, double* xt)
{
    return guard([&] {
        SG_CHECK(xt, SGIPC_EINVAL, "null output");
        ctx_of(h).get_xtilde(xt);
    });
}

int sgipc_get_kappa(sgipc_t h, double* k)
{
    return guard([&] {
        SG_CHECK(k, SGIPC_EINVAL, "null output");
        *k = ctx_of(h).kappa();
    });
}

int sgipc_set_kappa(sgipc_t h, double k)
{
    return guard([&] { ctx_of(h).set_kappa(k); });
}

int sgipc_set_state(sgipc_t h, const double* x, const double* v, const double* xt, double kappa)
{
    return guard([&] {
        SG_CHECK(x && v && xt, SGIPC_EINVAL, "null state array");
        ctx_of(h).set_state(x, v, xt, kappa);
    });
}

int sgipc_set_pcg(sgipc_t h, const sgipc_pcg_cfg* cfg)
{
    return guard([&] {
        SG_CHECK(cfg, SGIPC_EINVAL, "null config");
        ctx_of(h).set_pcg(*cfg);
    });
}

int sgipc_lbvh_build(sgipc_t h, int32_t which, uint64_t* keys, uint32_t* nodes, double* aabbs)
{
    return guard([&] {
        SG_CHECK(which == 0 || which == 1, SGIPC_EINVAL, "which must be 0 (faces) or 1 (edges)");
        ctx_of(h).lbvh_build(which, keys, nodes, aabbs);
    });
}

int sgipc_contact_query(sgipc_t h, int32_t* pairs, int32_t cap, int32_t* counts,
                        uint32_t* ground, int32_t gcap, int32_t* n_ground)
{
    return guard([&] { ctx_of(h).contact_query(pairs, cap, counts, ground, gcap, n_ground); });
}

int sgipc_assemble(sgipc_t h, double* grad, int32_t* nnzb, int32_t* rowptr, int32_t* cols,
                   double* vals, int32_t cap)
{
    return guard([&] { ctx_of(h).assemble_out(grad, nnzb, rowptr, cols, vals, cap); });
}

int sgipc_energy(sgipc_t h, double* terms)
{
    return guard([&] {
        SG_CHECK(terms, SGIPC_EINVAL, "null output");
        ctx_of(h).energy_terms(terms);
    });
}

int sgipc_set_soft_targets(sgipc_t ctx, const double* targets)
{
    return guard([&] {
        SG_CHECK(targets, SGIPC_EINVAL, "null targets");
        ctx_of(ctx).set_soft_targets(targets);
    });
}

int sgipc_get_animation_rate(sgipc_t ctx, double* rate)
{
    return guard([&] {
        SG_CHECK(rate, SGIPC_EINVAL, "null argument");
        *rate = ctx_of(ctx).animation_rate();
    });
}

int sgipc_set_dof_order(sgipc_t ctx, int32_t order)
{
    return guard([&] {
        SG_CHECK(order == SGIPC_DOF_FEM_FIRST || order == SGIPC_DOF_ABD_FIRST, SGIPC_EINVAL,
                 "unknown DOF order");
        ctx_of(ctx).set_dof_order(order);
    });
}

int sgipc_mm_export(sgipc_t ctx, const char* dir, int32_t frame, int32_t newton_step)
{
    return guard([&] {
        SG_CHECK(dir, SGIPC_EINVAL, "null directory");
        ctx_of(ctx).mm_export(dir, frame, newton_step);
    });
}

int sgipc_mm_export_step(sgipc_t ctx, const char* dir, int32_t frame)
{
    return guard([&] {
        SG_CHECK(dir, SGIPC_EINVAL, "null directory");
        ctx_of(ctx).mm_export_step(dir, frame);
    });
}

int sgipc_step_log(sgipc_t ctx, sgipc_iter_stats* out, int32_t cap, int32_t* n)
{
    return guard([&] {
        SG_CHECK(n, SGIPC_EINVAL, "null argument");
        const auto& v = ctx_of(ctx).step_log();
        *n            = (int32_t)v.size();
        if(out && cap > 0)
            std::copy(v.begin(), v.begin() + std::min<size_t>(v.size(), (size_t)cap), out);
    });
}

int sgipc_set_stream(sgipc_t ctx, void* stream)
{
    return guard([&] { ctx_of(ctx).set_stream(static_cast<hipStream_t>(stream)); });
}

int sgipc_ccd_bound(sgipc_t h, const double* dx, double* out)
{
    return guard([&] {
        SG_CHECK(dx && out, SGIPC_EINVAL, "null argument");
        ctx_of(h).ccd_bound(dx, out);
    });
}

int sgipc_ccd_candidates(sgipc_t h, const double* dx, double alpha, int32_t* pairs, int32_t cap,
                         int32_t* n)
{
    return guard([&] {
        SG_CHECK(dx && n && cap >= 0 && (pairs || cap == 0), SGIPC_EINVAL, "null argument");
        SG_CHECK(alpha >= 0.0 && alpha <= 1.0, SGIPC_EINVAL, "alpha must lie in [0, 1]");
        *n = ctx_of(h).ccd_candidates(dx, alpha, pairs, cap);
    });
}

int sgipc_newton_iteration(sgipc_t h, sgipc_iter_stats* st)
{
    return guard([&] { ctx_of(h).newton_iteration(st); });
}

int sgipc_time_step(sgipc_t h, sgipc_step_stats* st)
{
    return guard([&] { ctx_of(h).time_step(st); });
}

int sgipc_advance(sgipc_t h, int32_t n, sgipc_advance_stats* st)
{
    return guard([&] {
        SG_CHECK(n >= 0, SGIPC_EINVAL, "n must be non-negative");
        ctx_of(h).advance(n, st);
    });
}

int sgipc_profile_enable(sgipc_t h, int32_t every)
{
    return guard([&] { ctx_of(h).profile_enable(every); });
}

int sgipc_profile_query(sgipc_t h, const char* kernel, int32_t* launches, double* total_ms)
{
    return guard([&] {
        SG_CHECK(kernel, SGIPC_EINVAL, "null kernel name");
        ctx_of(h).profile_query(kernel, launches, total_ms);
    });
}

int sgipc_learned_mode(sgipc_t h, int32_t mode)
{
    return guard([&] { ctx_of(h).set_learned(mode); });
}

int sgipc_warm_start_adaptive(sgipc_t h, int32_t on)
{
    return guard([&] { ctx_of(h).set_warm_adaptive(on); });
}

int sgipc_learned_samples(sgipc_t h, float* feat, float* target, int32_t cap, int32_t* n)
{
    return guard([&] {
        SG_CHECK(feat && target && n && cap >= 0, SGIPC_EINVAL, "null output");
        *n = ctx_of(h).learned_samples(feat, target, cap);
    });
}

int sgipc_mlp_train(sgipc_t h, const float* feat, const float* target, int32_t n,
                    int32_t epochs, double lr, double* loss)
{
    return guard([&] {
        SG_CHECK(feat && target, SGIPC_EINVAL, "null batch");
        double L = ctx_of(h).mlp_train(feat, target, n, epochs, lr);
        if(loss)
            *loss = L;
    });
}

int sgipc_mlp_loss_grad(sgipc_t h, const float* feat, const float* target, int32_t n,
                        float* grad, double* loss)
{
    return guard([&] {
        SG_CHECK(feat && target, SGIPC_EINVAL, "null batch");
        double L = ctx_of(h).mlp_loss_grad(feat, target, n, grad);
        if(loss)
            *loss = L;
    });
}

int sgipc_mlp_set(sgipc_t h, const float* w)
{
    return guard([&] {
        SG_CHECK(w, SGIPC_EINVAL, "null weights");
        ctx_of(h).mlp_set(w);
    });
}

int sgipc_mlp_get(sgipc_t h, float* w)
{
    return guard([&] {
        SG_CHECK(w, SGIPC_EINVAL, "null weights");
        ctx_of(h).mlp_get(w);
    });
}

int sgipc_problem_sizes(sgipc_t h, int64_t* out)
{
    return guard([&] {
        SG_CHECK(out, SGIPC_EINVAL, "null output");
        ctx_of(h).problem_sizes(out);
    });
}

}  // extern "C"
