// context.h -- device-resident simulation state and the Newton driver
// (counterpart of the reference's GIPC class, StiffGIPC/GIPC.cuh:25-249).
#pragma once

#include "../../include/sgipc.h"
#include "learned.h"
#include "linsys.h"
#include "sim_kernels.h"

namespace sgipc
{

class Context
{
  public:
    explicit Context(const sgipc_scene_desc& d);
    ~Context();

    int  nv() const { return nv_; }
    void set_positions(const double* x);
    // state at a time-step boundary (x_prev = x), as the oracle's set_state
    void set_state(const double* x, const double* v, const double* xt, double kappa);
    void set_pcg(const sgipc_pcg_cfg& c);
    void get(const DBuf<double>& b, double* out);
    void set_xtilde(const double* xt);
    void get_xtilde(double* out) { get(xt_, out); }
    double kappa() const { return kappa_; }
    void   set_kappa(double k) { kappa_ = k; }
    // new positions of the animated soft targets ([3 n_soft])
    void   set_soft_targets(const double* t);
    // boundary DOF order (SGIPC_DOF_*) of assemble_out and mm_export
    void   set_dof_order(int order) { dof_order_ = order; }
    void   mm_export(const std::string& dir, int frame, int newton_step);
    // export the system of every solving Newton iteration of the next time
    // step (enable_matrix_market_export before IPC_Solver, main.cu:1105-1115)
    void   mm_export_step(const std::string& dir, int frame);
    // per-iteration statistics of the last time step (stats.json "newton")
    const std::vector<sgipc_iter_stats>& step_log() const { return step_log_; }
    // the caller's stream (nullptr: the context's own)
    void   set_stream(hipStream_t s);
    double animation_rate() const { return dev_.soft_rate; }

    // hot-path pieces exposed through the C ABI
    void lbvh_build(int which, uint64_t* keys, uint32_t* nodes, double* aabbs);
    void contact_query(int32_t* pairs, int cap, int32_t* counts, uint32_t* ground, int gcap,
                       int32_t* n_ground);
    void assemble_out(double* grad, int32_t* nnzb, int32_t* rowptr, int32_t* cols, double* vals,
                      int cap);
    void energy_terms(double* terms);
    // step-size bounds of a vertex direction dx (x_new = x - alpha dx) and the
    // full-CCD candidate set at a step (sgipc_ccd_bound / sgipc_ccd_candidates)
    void ccd_bound(const double* dx, double out[6]);
    int  ccd_candidates(const double* dx, double alpha, int32_t* out, int cap);
    void newton_iteration(sgipc_iter_stats* st);
    void time_step(sgipc_step_stats* st);
    void advance(int n, sgipc_advance_stats* st);
    void profile_enable(int every);
    void profile_query(const char* name, int32_t* launches, double* total_ms);
    void problem_sizes(int64_t* out);
    // learned coarse warm start: mode 0 off, 1 Galerkin-Jacobi guess, 2 network
    void set_learned(int mode);
    int  learned_samples(float* feat, float* target, int cap);
    void set_warm_adaptive(int on);
    double mlp_train(const float* X, const float* T, int n, int epochs, double lr)
    {
        return mlp_.train(X, T, n, epochs, lr, s_);
    }
    double mlp_loss_grad(const float* X, const float* T, int n, float* grad)
    {
        return mlp_.loss_grad(X, T, n, grad, s_);
    }
    void mlp_set(const float* w) { mlp_.set(w, s_); SG_HIP(hipStreamSynchronize(s_)); }
    void mlp_get(float* w) const { mlp_.get(w, s_); }

    DBuf<double> x_, v_;

  private:
    // reference phases
    void   build_bvh();                    // buildBVH
    void   build_cp();                     // buildCP (self + ground proximity)
    void   assemble(bool hessian);         // computeGradientAndHessian
    double energy();                       // computeEnergy
    bool   is_intersected();               // isIntersected
    double ground_bound(double slack);     // ground_largestFeasibleStepSize
    double self_bound(const int4* cand, const unsigned* count, int cap, double slack);
    int    full_ccd(double alpha);         // buildBVH_FULLCCD + buildFullCP
    double cfl_speed();
    void   step_bounds(double out[6]);
    void   line_search(double& alpha, double cfl_alpha, int* nls, double* e0, double* e1);
    void   suggest_kappa(double& k);
    void   upper_bound_kappa(double& k);
    void   init_kappa();
    void   check_overflow();
    void   begin_step();                   // IPC_Solver prologue (kappa)
    void   end_step();                     // updateVelocities + computeXTilta
    // after a Newton loop: true when the animation is complete and the step
    // ends; else the next loop at the next rate (IPC_Solver sub-steps)
    bool   next_newton_loop();
    double read_d(const double* p);
    void   step_forward(double alpha);
    void   body_fit(int b, const double* f, double* q) const;
    void   body_apply(int b, const double* q, double* f) const;
    unsigned read_u(const unsigned* p);

    hipStream_t s_ = nullptr, own_s_ = nullptr;
    int         dof_order_ = 0;
    // full BSR of the assembled system in the boundary DOF order
    void        gather_system(std::vector<double>& grad, std::vector<int>& rp,
                              std::vector<int>& cl, std::vector<double>& vl,
                              bool reassemble = true);
    std::vector<sgipc_iter_stats> step_log_;
    std::string mm_dir_;
    int         mm_frame_ = -1, mm_ns_ = 0;
    bool        mm_armed_ = false;
    sgipc_scene_desc desc_{};
    int nv_ = 0, nt_ = 0, nf_ = 0, ne_ = 0, ns_ = 0, ntri_ = 0, nh_ = 0;
    int nabd_ = 0, nfem_ = 0, nn_ = 0;  // affine bodies, FEM vertices, system nodes
    double dHat_ = 0, kappa_ = 0, bbox_diag2_ = 0, scene_diag_ = 0, mean_mass_ = 0;
    int    newton_k_ = 0;
    bool   in_step_  = false;
    Profiler prof_;
    double xdelta0_  = 0;

    DBuf<double>   xr_, xprev_, xt_, mass_, dx_, xsave_, gc_, gs_, rhs_, scratch_;
    DBuf<double>   dminv_, vol_, mu_, lam_;
    DBuf<int>      btype_, body_;
    DBuf<int>      gh_up_, gh_lo_, gh_ptr_, gh_src_, gv_ptr_, gv_src_;  // FEM gather lists
    DBuf<double>   fem_buf_;                                            // element records
    DBuf<uint32_t> tri_, hinge_;                                        // cloth
    DBuf<double>   tri_dminv_, tri_area_, hinge_th0_, hinge_len0_;
    DBuf<int>      bslots_, node_of_, bstart_, bfix_;                   // affine bodies
    DBuf<double>   xbar_, bM_, bV_, bq_, bqs_, gbody_, dy_;
    std::vector<int>    h_bstart_, h_bfix_;
    std::vector<double> h_xbar_, h_mass_, h_bMinv_, h_bG_;
    DBuf<int4>     fr_code_;                                            // lagged friction
    DBuf<double>   fr_lam_, fr_coord_, fr_basis_, fg_lam_;
    DBuf<uint32_t> fg_ids_;
    DBuf<unsigned> fr_count_;                                           // [0] self, [1] ground
    int            n_up_ = 0;
    DBuf<uint32_t> tets_, faces_, edges_, surf_, gp_;
    DBuf<unsigned> gcount_, pcount_, ccount_;
    DBuf<int4>     pairs_, pairs_raw_, cand_;
    DBuf<int>      overflow_;
    DBuf<double>   epart_, eterms_;
    std::vector<int> h_bmotor_;                                         // motor bodies
    DBuf<int>      bmotor_;
    DBuf<double>   bqm_;
    bool           any_motor_ = false;
    void           set_motor_targets();
    DBuf<uint32_t> soft_idx_;                                           // soft targets
    DBuf<double>   soft_tgt_;
    int            nsoft_ = 0;
    double         anim_sub_ = 1.0;  // animation_subRate = 1 / motion_rate
    int            pair_cap_ = 0, cand_cap_ = 0;

    Mlp          mlp_;
    Coarse       coarse_;
    int          learned_ = 0;
    // adaptive warm start: solves left to run without the guess, the current
    // pause length (doubles while the average stays negative), the probes
    // judged and their averaged net gain in ms
    int          warm_adaptive_ = 1, warm_skip_ = 0, warm_pause_ = 0, warm_probes_ = 0;
    double       warm_gain_     = 0.0;
    // PCG iterations of the recent zero-guess MAS solves (averaged with
    // weight 1/2): a warm-started solve is abandoned at twice this count
    double       plain_iters_   = 0.0;
    DBuf<double> x0_, b2_, x1_, b3_, zt_;
    PcgStats     solve_linear(const SysView& S, Precond pc, const PcgConfig& cfg, double* sol,
                              int* warm, int* warm_fallback);
    PcgStats     solve_plain(const SysView& S, Precond pc, const PcgConfig& cfg, double* sol);

    Bsr   A_;
    Coo   coo_;
    Mas   mas_;
    Pcg   pcg_;
    Lbvh  bf_, be_;
    SimDev dev_;
};

}  // namespace sgipc
