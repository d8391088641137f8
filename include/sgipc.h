/*
 * sgipc.h -- C ABI of libsgipc.so, the MI355X (gfx950) implementation of the
 * Stiff-GIPC Newton-step hot path.
 *
 * Conventions
 *   - every function returns 0 on success, a negative SGIPC_E* code on error;
 *     sgipc_last_error() returns a thread-local message for the last failure
 *     (the reference aborts instead: cuda_tools.h:172-206).
 *   - all pointers are HOST pointers; the library owns its device memory.
 *     Hot-path work runs on device-resident state; inputs are copied once.
 *   - 3x3 blocks are ROW-MAJOR doubles (reference: column-major Eigen 3x3,
 *     linear_system/linear_system/global_matrix.h:18-203).
 *   - BSR matrices at this boundary are full (both triangles) CSR-of-blocks:
 *     rowptr[nb+1], cols[nnzb], vals[9*nnzb].  Columns sorted within a row.
 *   - one context per host thread; calls on one context are not thread safe.
 */
#ifndef SGIPC_H
#define SGIPC_H

#include <stddef.h>
#include <stdint.h>
#include "sgipc_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SGIPC_OK          0
#define SGIPC_EINVAL     -1 /* bad argument / shape                 */
#define SGIPC_EHIP       -2 /* HIP runtime error                    */
#define SGIPC_ECAPACITY  -3 /* a device buffer capacity was exceeded */
#define SGIPC_ENUMERIC   -4 /* non-finite value / solver breakdown   */
#define SGIPC_ESTATE     -5 /* call out of order                     */

typedef struct sgipc_context* sgipc_t;
typedef struct sgipc_linear_system* sgipc_ls_t;

/* message for the last error on this thread ("" if none) */
const char* sgipc_last_error(void);
/* "gfx950 <name> CUs=<n>" of the current device, into buf */
int sgipc_device_info(char* buf, int32_t buflen);

/* ------------------------------------------------------------------------ */
/* PCG configuration and statistics                                          */
/* replaces gipc::PCGSolverConfig (linear_system/solver/pcg_solver.h:6-15)  */
/* ------------------------------------------------------------------------ */
typedef struct sgipc_pcg_cfg {
    double  rel_tol;          /* stop when rz <= rel_tol^2 * rz0 (or rr)       */
    double  abs_tol;          /* ... or rz <= abs_tol^2                        */
    int32_t max_iter;         /* 0 -> number of DOFs                           */
    int32_t use_preconditioned_norm; /* 1: rz test (reference default), 0: rr */
    int32_t check_every;      /* convergence test period (reference: 10)       */
    int32_t solver_precision; /* enum sgipc_solver_precision                   */
} sgipc_pcg_cfg;

typedef struct sgipc_pcg_stats {
    int32_t iterations;       /* PCG iterations (reference count semantics)    */
    int32_t outer_iterations; /* 1; 2 when the mixed solve fell back to FP64    */
    int32_t converged;
    int32_t pad;
    double  rel_residual;     /* ||b - A x||_2 / ||b||_2 recomputed in FP64    */
    double  rz0, rz;          /* preconditioned residual norms (last sweep)    */
    double  ms;               /* device time of the solve                      */
} sgipc_pcg_stats;

/* ------------------------------------------------------------------------ */
/* Stand-alone linear system: BSR + preconditioner + PCG                     */
/* replaces gipc::GlobalLinearSystem::solve_loaded_system                    */
/*   (global_linear_system.cu:407-490), PCGSolver::solve (pcg_solver.cu:155),*/
/*   MASPreconditioner::{setPreconditioner_bcoo, preconditioning}            */
/*   (MASPreconditioner.cu:2169-2251)                                        */
/* ------------------------------------------------------------------------ */

/* Create from a full symmetric BSR.  part_id (may be NULL) gives the MAS
 * 16-vertex bank of each block row (non-decreasing); MAS connectivity is the
 * off-diagonal block pattern.  precond: enum sgipc_precond. */
int sgipc_ls_create(int32_t nb, int32_t nnzb, const int32_t* rowptr,
                    const int32_t* cols, const double* vals,
                    const int32_t* part_id, int32_t precond,
                    int32_t mas_precision, sgipc_ls_t* out);
int sgipc_ls_destroy(sgipc_ls_t ls);
/* y = A x  (x, y: 3*nb doubles) */
int sgipc_ls_spmv(sgipc_ls_t ls, const double* x, double* y);
/* z = M^-1 r */
int sgipc_ls_precond_apply(sgipc_ls_t ls, const double* r, double* z);
/* solve A x = b from x0 = 0 */
int sgipc_ls_pcg(sgipc_ls_t ls, const double* b, double* x,
                 const sgipc_pcg_cfg* cfg, sgipc_pcg_stats* stats);
/* Create from the reference's merged upper BCOO (GIPCTripletMatrix after
 * Converter::convert, global_matrix.h:18-203, converter.cu:23-189): nnz
 * blocks (rows[k] <= cols[k], duplicates summed), values row-major 3x3 (the
 * reference stores column-major Eigen blocks: transpose them), every diagonal
 * block present.  Expanded to the full BSR of sgipc_ls_create. */
int sgipc_ls_create_bcoo(int32_t nb, int32_t nnz, const int32_t* rows,
                         const int32_t* cols, const double* vals,
                         const int32_t* part_id, int32_t precond,
                         int32_t mas_precision, sgipc_ls_t* out);
/* Solve-only mode (main.cu:387-646, :1047-1087): create from a MatrixMarket A
 * (coordinate real; general -- blocks below the diagonal dropped -- or
 * symmetric -- mirrored; square, rows divisible by 3, every diagonal 3x3
 * block present), read as to_upper_triangle_block_triplets does. */
int sgipc_ls_create_mm(const char* A_path, const int32_t* part_id, int32_t precond,
                       int32_t mas_precision, sgipc_ls_t* out);
/* the unique upper blocks a MatrixMarket A describes, read as
 * sgipc_ls_create_mm does (host only): *nb block rows, *nnz blocks; the first
 * min(nnz, cap) (row, col) pairs in (row, col) order and their row-major
 * values (9 doubles each) into rows/cols/vals (NULL with cap 0 to query) */
int sgipc_mm_read_upper_bcoo(const char* A_path, int32_t* nb, int32_t* nnz, int32_t* rows,
                             int32_t* cols, double* vals, int32_t cap);
/* block rows of a system */
int sgipc_ls_block_rows(sgipc_ls_t ls, int32_t* nb);
/* MatrixMarket array real M x 1 vector (read_mm_array_vector_mx1): *m = M;
 * the first min(M, cap) values into b (b may be NULL with cap 0) */
int sgipc_mm_read_vector(const char* path, double* b, int32_t cap, int32_t* m);
/* work of the system on the caller's stream from now on (hipStream_t;
 * NULL returns to the system's own stream) */
int sgipc_ls_set_stream(sgipc_ls_t ls, void* stream);

/* sgipc_ls_pcg without allocating.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The same solve (FP64 PCG
 * with the system's preconditioner, same kernels) with every work buffer --
 * device copies of b and x, PCG vectors, reductions, control words, the
 * block-Jacobi inverse -- in one fixed device array per GPU (about 31 MB,
 * held until the process ends).  Limits: at most SGIPC_NOALLOC_MAX_ROWS block
 * rows (SGIPC_ECAPACITY beyond), FP64 solver precision only (SGIPC_EINVAL
 * for the mixed mode), called with the system's device current.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU race.
 * stats->ms is host wall time of the call. */
#define SGIPC_NOALLOC_MAX_ROWS 65536
int sgipc_ls_pcg_noalloc(sgipc_ls_t ls, const double* b, double* x,
                         const sgipc_pcg_cfg* cfg, sgipc_pcg_stats* stats);
/* MAS hierarchy: levels, and level_sizes[2*(levels+1)] = (count, offset)
 * pairs in the reference's d_levelSize layout (MASPreconditioner.cu:1768). */
int sgipc_ls_mas_info(sgipc_ls_t ls, int32_t* levels, int32_t* level_sizes,
                      int32_t* total_clusters);
/* coarse node chain per vertex: chain[v*(levels-1) + l-1] = global id of
 * vertex v's node at level l (reference itable, _aggregationKernel:592) */
int sgipc_ls_mas_chain(sgipc_ls_t ls, int32_t* chain);
/* one bank's assembled 48x48 matrix mat (FP64, row-major, the Galerkin
 * product of its level restricted to the bank: the reference's MasMatrixSymT
 * unpacked, PrepareHessian_bcoo MASPreconditioner.cu:1810-2084) and the
 * inverse the apply uses, inv (__inverse6_P96x96 :628-724; widened to FP64
 * when the MAS is FP32).  Banks are numbered level by level (level 0 first)
 * as the clusters are (total_clusters / 16 banks); either output may be NULL */
int sgipc_ls_mas_bank(sgipc_ls_t ls, int32_t bank, double* mat, double* inv);

/* ------------------------------------------------------------------------ */
/* Spatial re-sorting keys (GIPC.cu:279-461 _calcTetMChash / _calcVertMChash /
 * hash_code, with the sort the reference runs on them).  kind 0: vertices,
 * the 30-bit Morton code of each position in the vertices' bounding box;
 * kind 1: tets, their centroid 0.25 ((a + b) + (c + d)) on the 1024^3 grid
 * packed in the axis order the box's extents select (hash_code types 0..5).
 * Keys are code << 32 | index, sorted ascending on the device; keys[n] and
 * order[n] (the index in each key's low word, n = n_verts or n_tets) may be
 * NULL.  x[3 n_verts] and tets[4 n_tets] are host arrays; a flat axis (zero
 * extent) contributes cell 0.                                                */
/* ------------------------------------------------------------------------ */
#define SGIPC_SORT_VERTICES 0
#define SGIPC_SORT_TETS 1
int sgipc_spatial_sort_keys(int32_t kind, const double* x, int32_t n_verts,
                            const uint32_t* tets, int32_t n_tets, uint64_t* keys,
                            uint32_t* order);

/* ------------------------------------------------------------------------ */
/* Simulation context (GIPC counterpart)                                     */
/* ------------------------------------------------------------------------ */
int sgipc_create(const sgipc_scene_desc* desc, sgipc_t* out);
int sgipc_destroy(sgipc_t ctx);

int sgipc_set_positions(sgipc_t ctx, const double* x);   /* [3N] */
int sgipc_get_positions(sgipc_t ctx, double* x);         /* [3N] */
int sgipc_get_velocities(sgipc_t ctx, double* v);        /* [3N] */
/* x_tilde = x_prev + dt v + dt^2 g on free vertices (GIPC.cu:8016-8038) */
int sgipc_set_xtilde(sgipc_t ctx, const double* xt);     /* [3N] */
int sgipc_get_xtilde(sgipc_t ctx, double* xt);           /* [3N] */
int sgipc_get_kappa(sgipc_t ctx, double* kappa);
int sgipc_set_kappa(sgipc_t ctx, double kappa);
/* Replace the whole integration state at a time-step boundary: positions
 * (also the step-start positions), velocities, x_tilde and kappa
 * (GIPC's vertexes / o_vertexes / velocities / xTilta / Kappa). */
int sgipc_set_state(sgipc_t ctx, const double* x, const double* v,
                    const double* xt, double kappa);
/* Replace the PCG settings of the Newton solve (PCGSolverConfig,
 * linear_system/solver/pcg_solver.h:6-15; wired at gipc/gipc.cu:52-57):
 * tolerances, stop norm and solver precision (FP64 or mixed). */
int sgipc_set_pcg(sgipc_t ctx, const sgipc_pcg_cfg* cfg);

/* LBVH over the surface faces (which = 0) or edges (which = 1) at the current
 * positions (lbvh_f/lbvh_e::Construct, mlbvh.cu:2074-2157).
 *   keys  [n]        sorted 64-bit (morton30 << 32 | index) keys
 *   nodes [4*(2n-1)] {parent, left, right, element} (0xFFFFFFFF = none)
 *   aabbs [6*(2n-1)] {upper xyz, lower xyz} per node
 * any output may be NULL. */
int sgipc_lbvh_build(sgipc_t ctx, int32_t which, uint64_t* keys,
                     uint32_t* nodes, double* aabbs);

/* Proximity contact set at the current positions (buildCP):
 *   pairs[4*cap] signed int4 codes (mlbvh.cu:563-1057 encoding),
 *   counts[5] = {total, unused, PP, PE, PT+EE} (reference h_cpNum),
 *   ground[gcap] vertex ids with (n.x - o)^2 <= dhat^2 (GIPC.cu:6435). */
int sgipc_contact_query(sgipc_t ctx, int32_t* pairs, int32_t cap,
                        int32_t* counts, uint32_t* ground, int32_t gcap,
                        int32_t* n_ground);

/* Gradient and Hessian at the current positions (computeGradientAndHessian,
 * GIPC.cu:10012-10228) on the system nodes (FEM vertices, then 4 nodes
 * (p, a1, a2, a3) per affine body; n_nodes from sgipc_problem_sizes).
 * grad[3 n_nodes] is the assembled right-hand side (zero on boundary
 * vertices and fixed bodies); rowptr[n_nodes + 1].  The Hessian is returned
 * as a full BSR; pass cap = 0 to query nnzb only. */
int sgipc_assemble(sgipc_t ctx, double* grad, int32_t* nnzb, int32_t* rowptr,
                   int32_t* cols, double* vals, int32_t cap);

/* Incremental potential at the current positions, FP64 (computeEnergy,
 * GIPC.cu:10392-10451).  terms[9] = {total, kinetic, elastic (tets),
 * barrier, ground, friction, cloth (membrane + bending), affine shape (and
 * the motor bodies' drive term, which the reference counts as kinetic),
 * soft targets}. */
int sgipc_energy(sgipc_t ctx, double* terms);

/* New positions of the animated soft targets, targets[3 n_soft] in the
 * order of sgipc_scene_desc.soft_index (the per-frame targetVert upload of
 * main.cu:1117-1140), and the current animation rate animation_fullRate
 * (0 before the first time step, 1 after a completed one). */
int sgipc_set_soft_targets(sgipc_t ctx, const double* targets);
int sgipc_get_animation_rate(sgipc_t ctx, double* rate);

/* DOF order of the system at this boundary (sgipc_assemble, sgipc_mm_export).
 * The library solves with FEM vertices first; the reference's
 * GlobalLinearSystem lays out [ABD bodies x 4 nodes | FEM vertices]
 * (global_linear_system.cu:263-336, subsystems created ABD first,
 * gipc/gipc.cu:42-47). */
#define SGIPC_DOF_FEM_FIRST 0
#define SGIPC_DOF_ABD_FIRST 1
int sgipc_set_dof_order(sgipc_t ctx, int32_t order);
/* MatrixMarket export of the Newton system at the current positions
 * (GlobalLinearSystem::export_matrix_market_files, global_linear_system.cu:
 * 98-261): <dir>/linear_system/frame_<frame, 5 digits>_ns_<newton_step>_A.mtx
 * (coordinate real general: every upper 3x3 block, then its transpose) and
 * ..._b.mtx (array real general, M x 1, the right-hand side of
 * sgipc_assemble), values %.17g, in the DOF order set above. */
int sgipc_mm_export(sgipc_t ctx, const char* dir, int32_t frame, int32_t newton_step);
/* export the system of every solving Newton iteration of the next time
 * step as sgipc_mm_export does, with newton_step counting from 0 (the
 * reference's export_linear_system_frames: enable_matrix_market_export
 * before IPC_Solver, main.cu:1105-1115, global_linear_system.cu:388-405) */
int sgipc_mm_export_step(sgipc_t ctx, const char* dir, int32_t frame);
/* work of the context on the caller's stream from now on (hipStream_t;
 * NULL returns to the context's own stream).  The reference runs on the
 * per-thread default stream. */
int sgipc_set_stream(sgipc_t ctx, void* stream);

/* Step-size bounds of a vertex direction dx [3 n_verts] at the current
 * positions (x_new = x - alpha dx; solve_subIP, GIPC.cu:10786-10822):
 * ground_largestFeasibleStepSize (GIPC.cu:8857-8915), ACCD
 * self_largestFeasibleStepSize over the proximity pairs (:8821-8855,
 * ACCD.cu:373-524), buildBVH_FULLCCD + buildFullCP candidates at that step
 * (mlbvh.cu:2093-2179) and the CFL bound (:8917-8968).
 * out[6] = {ground bound, self bound, full-CCD candidate count, CFL max
 * speed (0 without candidates), alpha_cfl, alpha (the bound the line search
 * starts from)}.  The contact set is the one of the current positions
 * (sgipc_contact_query's). */
int sgipc_ccd_bound(sgipc_t ctx, const double* dx, double* out);
/* full-CCD candidate codes (signed int4, a1 encoding) for the swept boxes of
 * x and x - alpha dx; *n = candidates found (may exceed cap) */
int sgipc_ccd_candidates(sgipc_t ctx, const double* dx, double alpha, int32_t* pairs,
                         int32_t cap, int32_t* n);

typedef struct sgipc_iter_stats {
    int32_t pcg_iterations;
    int32_t n_pairs;          /* self contact records                  */
    int32_t n_ground;         /* ground contacts                       */
    int32_t line_search_steps;
    int32_t converged;        /* Newton stop test met before solving   */
    int32_t solver_fallback;  /* 1: the mixed solve was redone in FP64 */
    int32_t warm_start;       /* solve continued from the affine (1) or the
                                 network-corrected (2) coarse guess        */
    int32_t warm_fallback;    /* 1: the warm-started solve was abandoned at
                                 twice the recent zero-guess PCG iteration
                                 count and solved again from zero          */
    double  pcg_rel_residual; /* FP64 ||b-Ax||/||b||                   */
    double  alpha;            /* accepted step                         */
    double  alpha_ccd;        /* CCD / ground bound before line search */
    double  xdelta;           /* ||dx||_inf of the previous direction  */
    double  energy0, energy1;
    double  ms_assembly, ms_solve, ms_ccd, ms_line_search, ms_total;
} sgipc_iter_stats;

/* One Newton iteration of solve_subIP (GIPC.cu:10746-10862): assemble,
 * stop test, MAS-PCG solve, ground/self CCD bound, line search, kappa. */
int sgipc_newton_iteration(sgipc_t ctx, sgipc_iter_stats* stats);

typedef struct sgipc_step_stats {
    int32_t newton_iterations;
    int32_t pcg_iterations;   /* summed over the step */
    double  ms_total;
    int32_t capped;           /* 1: the step ended at the Newton iteration cap
                                 (newton_max_iter; reference iterCap = 10000,
                                 GIPC.cu:10739) without meeting the stop test */
    int32_t pad;
} sgipc_step_stats;

/* One time step (IPC_Solver, GIPC.cu:10940-11131): kappa setup, Newton loop,
 * velocity update, next x_tilde. */
int sgipc_time_step(sgipc_t ctx, sgipc_step_stats* stats);
/* statistics of every Newton iteration of the last time step (the stats.json
 * "newton" list, main.cu:1144-1150), the final stop-test pass of each Newton
 * loop included (converged = 1); *n = count, the first min(n, cap) to out */
int sgipc_step_log(sgipc_t ctx, sgipc_iter_stats* out, int32_t cap, int32_t* n);

typedef struct sgipc_advance_stats {
    int32_t newton_iterations; /* Newton solves performed (== n requested)  */
    int32_t time_steps_begun;
    int32_t pcg_iterations;
    int32_t solver_fallbacks;  /* mixed-mode solves redone in FP64          */
    double  max_pcg_rel_residual;
    double  ms_total;
    int32_t warm_starts;       /* solves continued from a coarse guess      */
    int32_t capped_steps;      /* time steps ended at the Newton cap without
                                  meeting the stop test                      */
    /* device time of the counted iterations by phase (HIP events): assembly
     * + stop test, MAS build + solve, CCD bounds, line search (with the next
     * contact detection) */
    double  ms_assembly, ms_solve, ms_ccd, ms_line_search;
    int32_t line_search_steps; /* Armijo backtracking steps, summed         */
    int32_t ccd_limited;       /* iterations whose CCD/CFL bound was < 1     */
    int32_t warm_fallbacks;    /* warm-started solves abandoned (see
                                  sgipc_iter_stats.warm_fallback)           */
    int32_t pad;
} sgipc_advance_stats;

/* Advance the integration by exactly n Newton iterations that solve (the
 * stop test's final assemble-and-check pass of a time step is not counted),
 * ending and beginning time steps as IPC_Solver does (GIPC.cu:10940-11131).
 * Used to time the hot path per Newton iteration. */
int sgipc_advance(sgipc_t ctx, int32_t n, sgipc_advance_stats* stats);

/* Per-kernel device timing with HIP events on the library stream.
 * Names: "spmv" (PCG A p), "mas_apply_l0", "fem", "mas_invert".
 * every >= 1 clears and starts collection, timing every every-th launch of
 * each name (an event pair costs microseconds of stream time); 0 stops it.
 * query returns the number of timed launches and their summed duration. */
int sgipc_profile_enable(sgipc_t ctx, int32_t every);
int sgipc_profile_query(sgipc_t ctx, const char* kernel, int32_t* launches,
                        double* total_ms);
/* ---- learned affine coarse warm start (no counterpart in the reference) ----
 * MAS's coarse spaces hold one translation per aggregate.  Each aggregate of
 * the coarsest MAS level (its free FEM nodes) gets an affine space
 * x_i = p + A (x_i - centroid) = J_i q (q in R^12, as for affine bodies):
 * q_c = (P_c^T A P_c)^-1 P_c^T b per aggregate, then, in mode 2,
 * q_c += |q_c| MLP(q_c / |q_c|, r_c / |r_c|) with r_c = P_c^T b (layers
 * 24 -> 128 -> 128 -> 12 with ReLU; BF16 weights, FP32 accumulation).  The
 * PCG solve continues from x0 = P q with the zero-guess stop threshold; plain
 * MAS judges the guess, and the zero-guess solve is the fallback. */
#define SGIPC_MLP_PARAMS 21260 /* W1[24][128] b1[128] W2[128][128] b2[128] W3[128][12] b3[12] */
#define SGIPC_MLP_IN 24
#define SGIPC_MLP_OUT 12
/* mode 0 off, 1 affine Galerkin guess only, 2 guess corrected by the network */
int sgipc_learned_mode(sgipc_t ctx, int32_t mode);
/* on (the default): each warm start's net gain is its predicted saving (from
 * the guess's preconditioned residual and PCG's linear convergence) minus its
 * own device time; while the running average (weight 1/2) of the net gains is
 * negative, the guesses are skipped for the next 2 solves (8 when the probe
 * saved less than a quarter of its cost), the pause doubling up to 32; off:
 * every solve computes and judges the guesses */
int sgipc_warm_start_adaptive(sgipc_t ctx, int32_t on);
/* training samples from the last solving Newton iteration: features
 * [n][SGIPC_MLP_IN] and targets [n][SGIPC_MLP_OUT] ((q* - q_c) / |q_c| with q*
 * the least-squares affine fit of the solution on the aggregate) */
int sgipc_learned_samples(sgipc_t ctx, float* feat, float* target, int32_t cap, int32_t* n);
/* full-batch Adam on the GPU (FP32 master weights, BF16 forward); final loss */
int sgipc_mlp_train(sgipc_t ctx, const float* feat, const float* target, int32_t n,
                    int32_t epochs, double lr, double* loss);
/* mean squared error and its gradient (SGIPC_MLP_PARAMS floats) */
int sgipc_mlp_loss_grad(sgipc_t ctx, const float* feat, const float* target, int32_t n,
                        float* grad, double* loss);
int sgipc_mlp_set(sgipc_t ctx, const float* weights);
int sgipc_mlp_get(sgipc_t ctx, float* weights);

/* problem sizes for algorithmic byte counts: out[8] = {n_nodes (system block
 * rows: FEM vertices + 4 per affine body), n_tets (elastic), nnzb static,
 * coo blocks, mas banks, mas levels, self pairs, ground pairs} */
int sgipc_problem_sizes(sgipc_t ctx, int64_t* out);

#ifdef __cplusplus
}
#endif
#endif /* SGIPC_H */
