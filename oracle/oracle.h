/*
 * oracle.h -- CPU FP64 restatement of the Stiff-GIPC Newton-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for libsgipc.so:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  Nothing in the product path links or calls it.
 *
 * Every function cites the reference file:line whose algorithm it restates.
 * Matrices are row-major; 3x3 blocks are row-major 9-double arrays.
 * Built with -ffp-contract=off so FP64 results do not depend on FMA fusion.
 */
#ifndef SGIPC_ORACLE_H
#define SGIPC_ORACLE_H

#include <stdint.h>
#include "../include/sgipc_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- dense helpers (orc_math.c) ---------------- */
/* cyclic Jacobi eigen-decomposition of a symmetric n x n matrix A (row-major).
 * w[n] ascending eigenvalues, V[n*n] with eigenvector k in column k. */
void orc_sym_eig(int n, const double* A, double* w, double* V);
/* A <- V max(w,0) V^T  (PSD projection by eigenvalue clamping) */
void orc_psd_project(int n, double* A);
/* in-place inverse by Gauss-Jordan with partial pivoting; returns 0 if ok */
int orc_inverse(int n, double* A);
/* in-place Gauss-Jordan inverse without pivoting, in the operation order of
 * the reference's MAS bank inverse (MASPreconditioner.cu:628-724) */
void orc_inverse_gj(int n, double* M);
/* host threads of the parallel loops; results do not depend on the count */
void orc_set_threads(int n);
int  orc_get_threads(void);

/* ---------------- FEM (orc_fem.c) ---------------- */
double orc_snk1_psi(const double F[9], double mu, double lam);
void   orc_snk1_pk1(const double F[9], double mu, double lam, double P[9]);
/* 9x9 Hessian d2psi/dvec(F)^2, vec column-major (index i + 3 j); project=1
 * clamps negative eigenvalues */
void   orc_snk1_hess(const double F[9], double mu, double lam, int project,
                     double H[81]);
/* Dm^-1, |volume|, Lame mu, lambda, lumped mass (initFEM, main.cu:310-356) */
void   orc_fem_precompute(int nv, const double* x_rest, int nt,
                          const uint32_t* tets, const double* E,
                          const double* nu, const double* rho, double* dminv,
                          double* vol, double* mu, double* lam, double* mass);
/* per-tet dt^2 vol-scaled gradient (12) and projected Hessian (12x12)
 * (_calculate_fem_gradient_hessian, femEnergy.cu:1789-1926) */
void   orc_tet_grad_hess(const double* x, const uint32_t t[4],
                         const double dminv[9], double vol, double mu,
                         double lam, double dt, double g[12], double H[144]);
double orc_tet_energy(const double* x, const uint32_t t[4],
                      const double dminv[9], double vol, double mu, double lam);

/* ground IPC barrier b(d^2) = -(d^2 - dhat2)^2 ln(d^2/dhat2) per vertex,
 * unscaled by kappa (GIPC.cu:6465-6516, 6839-6860) */
double orc_ground_energy(double dist2, double dhat2);
/* gradient w.r.t. x (3) and 3x3 block, scaled by kappa */
void   orc_ground_grad_hess(const double x[3], const double n[3], double off,
                            double dhat2, double kappa, double g[3],
                            double H[9]);

/* ---------------- LBVH (orc_lbvh.c) ---------------- */
uint32_t orc_expand_bits(uint32_t v);
uint32_t orc_morton_code(double x, double y, double z);
/* orc_sort.c: spatial re-sorting keys (kind 0 vertices, 1 tets) */
uint32_t orc_hash_code(int type, double x, double y, double z);
int      orc_tet_hash_type(double sx, double sy, double sz);
int      orc_spatial_keys(int kind, const double* x, int nv, const uint32_t* tets, int nt,
                          uint64_t* keys, uint32_t* order);
/* leaf AABBs {upper xyz, lower xyz} of n primitives of arity 2 or 3 */
void orc_leaf_aabbs(const double* x, const uint32_t* prims, int arity, int n,
                    double* aabb);
/* full construction: keys[n] sorted, nodes[4*(2n-1)], aabbs[6*(2n-1)]
 * (lbvh_f::Construct, mlbvh.cu:2074-2091; Karras 2012) */
void orc_lbvh_build(const double* x, const uint32_t* prims, int arity, int n,
                    uint64_t* keys, uint32_t* nodes, double* aabbs);
/* same from precomputed leaf boxes (used for swept CCD boxes) */
void orc_lbvh_build_boxes(const double* leaf, int n, uint64_t* keys,
                          uint32_t* nodes, double* aabbs);

/* ---------------- contacts (orc_contact.c) ---------------- */
double orc_d_pp(const double* v0, const double* v1);
double orc_d_pe(const double* v0, const double* v1, const double* v2);
double orc_d_pt(const double* v0, const double* v1, const double* v2,
                const double* v3);
double orc_d_ee(const double* v0, const double* v1, const double* v2,
                const double* v3);
int    orc_dtype_pt(const double* v0, const double* v1, const double* v2,
                    const double* v3);
int    orc_dtype_ee(const double* v0, const double* v1, const double* v2,
                    const double* v3);
/* proximity pairs by brute force over (surface vertex, face) and
 * (edge, edge); returns the number written (<= cap); counts[5] as the
 * reference's h_cpNum.  pairs are the signed int4 codes. */
int orc_contact_pairs_bruteforce(const sgipc_scene_desc* s, const double* x,
                                 double dhat2, int32_t* pairs, int cap,
                                 int32_t counts[5]);
/* same set via the LBVH traversal (mlbvh.cu:1308-1634) */
int orc_contact_pairs_bvh(const sgipc_scene_desc* s, const double* x,
                          double dhat2, int32_t* pairs, int cap,
                          int32_t counts[5]);
/* barrier energy of one contact code (kappa included, GIPC.cu:703-959) */
double orc_barrier_energy(const double* x, const double* x_rest,
                          const int32_t code[4], double dhat2, double kappa);
/* gradient (on up to 4 vertices) and GIPC Gauss-Newton Hessian; returns the
 * number of vertices nv (2, 3 or 4); ids[nv]; g[3*nv]; H[(3nv)^2]
 * (_calBarrierGradientAndHessian, GIPC.cu:3108-4947) */
int orc_barrier_grad_hess(const double* x, const double* x_rest,
                          const int32_t code[4], double dhat2, double kappa,
                          int32_t ids[4], double* g, double* H);
/* second derivative of kappa dhat2^2 (I-1)^2 ln^2 I in s = sqrt(I), with the
 * small-distance clamp (GIPC.cu:3272-3289) */
double orc_barrier_lambda0(double I5, double dhat2, double kappa);

/* ---------------- linear system (orc_linsys.c) ---------------- */
typedef struct orc_bsr {
    int nb, nnzb;
    int* rowptr; /* nb+1 */
    int* cols;   /* nnzb */
    double* vals;/* 9*nnzb row-major blocks */
    /* the last 4 nbody block rows are affine bodies (p, a1, a2, a3): the
     * block-Jacobi preconditioner inverts each body's 12x12 diagonal block
     * (ABDPreconditioner) */
    int nbody;
} orc_bsr;

/* merge upper triplets (row <= col) into a full symmetric BSR
 * (Converter::convert, converter.cu:23-189) */
int  orc_bsr_from_upper_triplets(int nb, int ntrip, const int* rows,
                                 const int* cols, const double* vals,
                                 orc_bsr* out);
void orc_bsr_free(orc_bsr* A);
void orc_bsr_spmv(const orc_bsr* A, const double* x, double* y);

typedef struct orc_mas orc_mas;
/* build the MAS hierarchy + block inverses for A (FEM range = all rows).
 * part_id: bank ids (non-decreasing, NULL -> runs of 16).  fp32 = 1 rounds
 * inverse blocks and R/Z to float as the reference (eigen_data.h:125). */
orc_mas* orc_mas_create(const orc_bsr* A, const int32_t* part_id, int fp32);
void     orc_mas_free(orc_mas* m);
void     orc_mas_apply(const orc_mas* m, const double* r, double* z);
int      orc_mas_levels(const orc_mas* m);
void     orc_mas_level_sizes(const orc_mas* m, int32_t* sizes);
int      orc_mas_total_clusters(const orc_mas* m);
void     orc_mas_chain(const orc_mas* m, int32_t* chain);

/* block-Jacobi preconditioner: 3x3 node blocks, 12x12 body blocks */
void orc_block_jacobi(const orc_bsr* A, const double* r, double* z);
/* PCG (pcg_solver.cu:171-304).  precond: 0 block-Jacobi, 1 MAS, -1 none */
int orc_pcg(const orc_bsr* A, const orc_mas* mas, int precond,
            const double* b, double* x, double rel_tol, double abs_tol,
            int max_iter, int use_rz, int check_every, double* rz0_out,
            double* rz_out);

/* ---------------- lagged friction (orc_friction.c) ---------------- */
/* lagged normal force lam, closest-point coordinates and 3x2 row-major
 * tangent basis of a contact code at x; returns the kind (0 PP, 1 PE, 2 PT,
 * 3 EE) or -1 when the code carries no friction */
int    orc_friction_lag(const double* x, const int32_t* code, double dhat2,
                        double kappa, double* lam, double coord[2],
                        double basis[6]);
double orc_friction_lag_ground(const double* x, const double n[3], double off,
                               double dhat2, double kappa);
/* unscaled energy lam f0(|u|) of one lagged pair (u: tangential relative
 * displacement since xp) */
double orc_friction_energy(const double* x, const double* xp,
                           const int32_t* code, const double coord[2],
                           const double basis[6], double lam, double eps2);
/* coef-scaled gradient (3m) and PSD Hessian (3m x 3m, row-major); returns m */
int    orc_friction_grad_hess(const double* x, const double* xp,
                              const int32_t* code, const double coord[2],
                              const double basis[6], double lam, double eps2,
                              double coef, int32_t ids[4], double g[12],
                              double H[144]);
double orc_friction_ground_energy(const double* x, const double* xp,
                                  const double n[3], double lam, double eps2);
void   orc_friction_ground_grad_hess(const double* x, const double* xp,
                                     const double n[3], double lam, double eps2,
                                     double coef, double g[3], double H[9]);

/* ---------------- cloth (orc_cloth.c) ---------------- */
/* stretch, shear, bend stiffnesses from the scene's cloth moduli
 * (main.cu:303-307) */
void   orc_cloth_params(double E, double nu, double bend_E, double thickness,
                        double strain_rate, double* stretch, double* shear,
                        double* bend);
/* rest-frame 2x2 Dm^-1 (row-major) and rest area of a triangle
 * (__calculateDm2D_double, femEnergy.cu:262-309) */
void   orc_tri_precompute(const double* x_rest, const uint32_t t[3],
                          double dminv[4], double* area);
/* Baraff-Witkin energy density, PK1 (3x2 row-major) and 6x6 vec(F)
 * Hessian (femEnergy.cu:360-412, 565-740, 1003-1041) */
double orc_bw_psi(const double F[6], double stretch, double shear, double r);
void   orc_bw_pk1(const double F[6], double stretch, double shear, double r,
                  double P[6]);
void   orc_bw_hess(const double F[6], double stretch, double shear, double r,
                   int project, double H[36]);
/* area-scaled (no dt^2) energy; dt^2 area-scaled gradient (9) and Hessian
 * (9x9) (_calculate_triangle_fem_gradient_hessian, femEnergy.cu:2215-2300) */
double orc_tri_energy(const double* x, const uint32_t t[3],
                      const double dminv[4], double area, double stretch,
                      double shear, double r);
void   orc_tri_grad_hess(const double* x, const uint32_t t[3],
                         const double dminv[4], double area, double stretch,
                         double shear, double r, double dt, double g[9],
                         double H[81]);
/* signed dihedral angle of hinge (q0, q1 | q2, q3) with its gradient (12)
 * and Hessian (12x12) when non-NULL (LibShell::edgeTheta,
 * femEnergy.cu:142-260) */
double orc_dihedral(const double* q0, const double* q1, const double* q2,
                    const double* q3, double g[12], double H[144]);
/* k |e0| (theta - theta0)^2 (no dt^2) and the dt^2-scaled gradient and PSD
 * Hessian (femEnergy.cu:2431-2566, 2590-2630) */
double orc_bend_energy(const double* x, const uint32_t hinge[4], double theta0,
                       double len0, double bend);
void   orc_bend_grad_hess(const double* x, const uint32_t hinge[4],
                          double theta0, double len0, double bend, double dt,
                          double g[12], double H[144]);
/* interior edges of a triangle list as hinges {a, b, opp0, opp1}, a < b,
 * in increasing (a, b) order (load_mesh.cpp:320-379); returns the count and
 * writes hinges when non-NULL */
int    orc_bend_hinges(int ntri, const uint32_t* tris, uint32_t* hinges);

/* ---------------- affine bodies (orc_abd.c) ---------------- */
/* ||A A^T - I||_F^2 of q = (p, a1, a2, a3), its gradient over (a1, a2, a3)
 * and unprojected 9x9 Hessian (details/abd_energy.inl:9-127) */
double orc_abd_shape_energy(const double q[12]);
void   orc_abd_shape_grad(const double q[12], double g[9]);
void   orc_abd_shape_hess(const double q[12], double H[81]);
/* x = p + A xbar; J^T v (12); J (3x12 row-major) (abd_jacobi_matrix.h:8-62) */
void   orc_abd_x(const double xbar[3], const double q[12], double x[3]);
void   orc_abd_jt(const double xbar[3], const double v[3], double out[12]);
void   orc_abd_jmat(const double xbar[3], double J[36]);
/* motor target q_p of a body from its prediction q~ and the angle speed * dt */
void   orc_abd_motor_target(const double qt[12], double theta, double qp[12]);

/* ---------------- learned warm start network (orc_mlp.c) ---------------- */
float  orc_bf16_round(float f);
int    orc_mlp_params(void);
/* loss (mean squared error) and, when grad != NULL, its gradient (FP64) at
 * FP32 master weights w (matrices used rounded to BF16); Y[n][12] optional */
double orc_mlp_loss_grad(const float* w, const float* X, const float* T, int n,
                         double* grad, float* Y);

/* ---------------- world / Newton (orc_world.c) ---------------- */
typedef struct orc_world orc_world;
orc_world* orc_world_create(const sgipc_scene_desc* s);
void       orc_world_free(orc_world* w);
void       orc_world_set_positions(orc_world* w, const double* x);
void       orc_world_set_positions_exact(orc_world* w, const double* x);
void       orc_world_set_soft_targets(orc_world* w, const double* t);
void       orc_world_set_anim_rate(orc_world* w, double r);
double     orc_world_get_anim_rate(const orc_world* w);
void       orc_world_get_positions(const orc_world* w, double* x);
void       orc_world_set_xtilde(orc_world* w, const double* xt);
void       orc_world_get_xtilde(const orc_world* w, double* xt);
double     orc_world_get_kappa(const orc_world* w);
void       orc_world_set_kappa(orc_world* w, double k);
/* system nodes: FEM vertices, then 4 per affine body (p, a1, a2, a3) */
int  orc_world_nodes(const orc_world* w);
/* assembled RHS gradient (3 nodes) and full node BSR Hessian (caller frees) */
int  orc_world_assemble(orc_world* w, double* grad, orc_bsr* H);
/* terms[8] as sgipc_energy */
/* terms[9] = {total, kinetic, elastic, barrier, ground, friction, cloth,
 * affine shape, soft targets} */
void orc_world_energy(orc_world* w, double* terms);
/* one Newton iteration; stats fields mirror sgipc_iter_stats */
int  orc_world_newton_iteration(orc_world* w, double* alpha_out,
                                int* pcg_iters_out, double* xdelta_out,
                                int* converged_out);
int  orc_world_time_step(orc_world* w, int* newton_iters_out,
                         int* pcg_iters_out);
/* state at a step boundary (x = x_prev) and kappa; rebuilds contacts */
void orc_world_set_state(orc_world* w, const double* x, const double* v,
                         const double* xt, double kappa);
/* exactly n solving Newton iterations across step boundaries */
int  orc_world_advance(orc_world* w, int n, int* steps_out, int* pcg_out);
/* step-size bounds of a vertex direction dx [3 nv] at the current positions
 * (solve_subIP, GIPC.cu:10786-10822): out[6] = {ground bound, ACCD bound over
 * the proximity pairs, full-CCD candidate count, CFL max speed (0 without
 * candidates), alpha_cfl, resulting alpha} */
void orc_world_ccd_bound(orc_world* w, const double* dx, double* out);
/* full-CCD candidate codes for x and x - alpha dx (mlbvh.cu:2093-2179) */
int  orc_world_ccd_candidates(orc_world* w, const double* dx, double alpha,
                              int32_t* out, int cap);
/* second derivative of the barrier in squared distance (GIPC.cu:9746-9750) */
double orc_barrier_H_b(double d, double dhat);
/* out[5] = {suggestKappa, upperBoundKappa's cap, mean mass, bbox diag^2,
 * dHat (squared)} (GIPC.cu:9752-9786) */
void   orc_world_kappa_info(orc_world* w, double* out);
/* initKappa (GIPC.cu:9789-9831) from kappa = 0 on the current contact set */
double orc_world_init_kappa(orc_world* w);
/* additive CCD time of impact (ACCD.cu:373-524), eta = 1 - slackness */
double orc_pt_ccd(const double* p, const double* t0, const double* t1,
                  const double* t2, const double* dp, const double* dt0,
                  const double* dt1, const double* dt2, double eta,
                  double thickness);
double orc_ee_ccd(const double* a0, const double* a1, const double* b0,
                  const double* b1, const double* da0, const double* da1,
                  const double* db0, const double* db1, double eta,
                  double thickness);

#ifdef __cplusplus
}
#endif
#endif
