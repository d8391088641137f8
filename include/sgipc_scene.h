/*
 * sgipc_scene.h -- plain-data scene descriptor shared by the HIP library
 * (libsgipc.so) and the CPU oracle.  Host arrays only; the library copies
 * everything it needs to device memory at sgipc_create().
 *
 * Mirrors the inputs the reference gathers in StiffGIPC/main.cu:263-356
 * (apply_settings, initFEM) and StiffGIPC/main.cu:47-75 (setMAS_partition),
 * flattened to C arrays:
 *   - positions are interleaved xyz doubles (reference: double3 AoS),
 *   - tets are 4 x uint32 (reference: uint4), faces 3 x uint32, edges 2 x uint32,
 *   - dhat is an absolute DISTANCE in metres; the squared threshold used by
 *     every contact test is dhat*dhat (reference: dHat = rel^2 * bboxDiag^2,
 *     GIPC.cu:8543, already squared).
 */
#ifndef SGIPC_SCENE_H
#define SGIPC_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* preconditioner choice (reference: pcg_data.P_type, gipc/gipc.cu:60-72) */
enum sgipc_precond {
    SGIPC_PRECOND_DIAG = 0, /* 3x3 block Jacobi (DiagPreconditioner)          */
    SGIPC_PRECOND_MAS  = 1  /* multilevel additive Schwarz (MAS_Preconditioner) */
};

/* arithmetic of the MAS block inverses and multilevel R/Z vectors */
enum sgipc_mas_precision {
    SGIPC_MAS_FP32 = 0, /* reference-faithful: FP64 assembly, FP32 inverse/apply */
    SGIPC_MAS_FP64 = 1  /* all-FP64 MAS                                           */
};

/* linear-solver precision mode.  MIXED: when the FP32 iteration stalls or
 * breaks down, FP64 PCG continues from its iterate, and the next 2 mixed
 * solves of the same context or system run in FP64 directly (the pause
 * doubling up to 32 while FP32 attempts keep failing; restarted when the PCG
 * settings change) */
enum sgipc_solver_precision {
    SGIPC_SOLVE_FP64  = 0, /* FP64 SpMV / PCG vectors (reference default)          */
    SGIPC_SOLVE_MIXED = 1  /* FP32 SpMV+MAS inner PCG, FP64 residual refinement    */
};

typedef struct sgipc_scene_desc {
    /* ---- vertices ---- */
    int32_t         n_verts;
    const double*   x;       /* [3*n_verts] positions at the start              */
    const double*   x_rest;  /* [3*n_verts] rest positions (NULL -> x)          */
    const double*   v;       /* [3*n_verts] velocities (NULL -> zero)           */
    const int32_t*  btype;   /* [n_verts] 0 free, !=0 fixed/obstacle (NULL -> 0)*/
    const int32_t*  body_id; /* [n_verts] body id, -1 = self contact allowed    */
    const int32_t*  part_id; /* [n_verts] 16-vertex bank id, non-decreasing in
                                vertex order (NULL -> runs of 16 vertices)       */
    /* ---- volumetric elements ---- */
    int32_t         n_tets;
    const uint32_t* tets;    /* [4*n_tets]                                      */
    const double*   tet_E;   /* [n_tets] Young's modulus                        */
    const double*   tet_nu;  /* [n_tets] Poisson ratio                          */
    const double*   tet_rho; /* [n_tets] density                                */
    /* ---- cloth (triangle shells; reference scene JSON "triangle_*" keys) ---- */
    int32_t         n_tris;
    const uint32_t* tris;    /* [3*n_tris] cloth triangles; their interior edges
                                carry the bending hinges                         */
    double          cloth_E;           /* triangle_mesh_youngs_modulus          */
    double          cloth_nu;          /* poisson_rate used for the cloth       */
    double          cloth_bend_E;      /* triangle_bend_youngs_modulus          */
    double          cloth_thickness;   /* triangle_mesh_thickness               */
    double          cloth_density;     /* triangle_mesh_density                 */
    double          cloth_strain_rate; /* strain_rate (cubic stretch stiffening)*/
    /* ---- affine bodies (ABD; reference abd_system/) ----
     * A vertex with abd_body[v] = b >= 0 moves with affine body b:
     * x = p_b + A_b (x0 - c_b), c_b the body's mass centre at creation.
     * ABD vertices come after every other vertex, grouped by body in
     * increasing body order; a tet is either all-ABD (one body: it only
     * supplies mass and volume) or all-FEM.  A body is fixed when its
     * vertices have btype != 0.                                            */
    int32_t         n_abd;
    const int32_t*  abd_body;  /* [n_verts] body index or -1 (NULL -> no ABD)  */
    double          abd_kappa; /* shape stiffness kappa (reference 1e8)        */
    /* ---- collision surface ---- */
    int32_t         n_faces;
    const uint32_t* faces;      /* [3*n_faces] */
    int32_t         n_edges;
    const uint32_t* edges;      /* [2*n_edges] */
    int32_t         n_surf_verts;
    const uint32_t* surf_verts; /* [n_surf_verts] */
    /* ---- physics ---- */
    double  dt;
    double  dhat;           /* contact distance threshold (m), squared internally */
    double  gravity[3];
    int32_t use_ground;
    double  ground_normal[3];
    double  ground_offset;
    double  kappa;          /* barrier stiffness; <= 0 -> adaptive (suggestKappa) */
    double  friction_mu;    /* self-contact friction coefficient (friction_rate)   */
    double  ground_friction_mu; /* ground friction coefficient (gd_friction_rate)   */
    /* ---- Newton / line search (reference JSON "settings") ---- */
    double  abs_xdelta_tol; /* stop when |dx|_inf < scene_diag * abs_xdelta_tol  */
    double  rel_xdelta_tol; /* ... or < rel_xdelta_tol * |dx0|_inf               */
    double  armijo_c1;
    double  armijo_beta;
    double  armijo_alpha_min;
    int32_t newton_max_iter;
    /* ---- linear solver ---- */
    double  pcg_rel_tol;
    double  pcg_abs_tol;
    int32_t pcg_max_iter;   /* 0 -> number of DOFs                              */
    int32_t pcg_use_preconditioned_norm;
    int32_t precond;        /* enum sgipc_precond                               */
    int32_t mas_precision;  /* enum sgipc_mas_precision                         */
    int32_t solver_precision; /* enum sgipc_solver_precision                    */
    /* ---- animated soft targets (reference TetMesh.targetIndex / targetVert,
     * scene JSON "animation" + settings motion_stiffness / motion_rate;
     * GIPC.cu:6365-6435, 7527-7560, 10941-11060) ----
     * vertex soft_index[i] is pulled towards soft_target[i] by the energy
     * 0.5 k |r (x - t)|^2, k = soft_stiffness, r the animation rate: a time
     * step runs Newton loops at r = s, 2 s, ... until r > 0.99, with
     * s = 1 / motion_rate (animation_subRate), the lagged friction sets rebuilt
     * between loops.  Free FEM vertices only, each at most once.            */
    int32_t         n_soft;
    const uint32_t* soft_index;   /* [n_soft]                                 */
    const double*   soft_target;  /* [3*n_soft] (sgipc_set_soft_targets moves them) */
    double          soft_stiffness; /* motion_stiffness (softMotionRate)      */
    double          motion_rate;    /* <= 0 -> 1                               */
    /* ---- scene extents (GIPC::init / main.cu:853-863) ----
     * bbox_diag: sqrt(bboxDiagSize2), the diagonal of the largest-volume
     * object box that scales kappa and the friction epsilon (<= 0: the AABB
     * of the surface faces); scene_diag: the diagonal of the union box of the
     * non-obstacle objects that scales abs_xdelta_tol (<= 0: the AABB of the
     * free vertices) */
    double          bbox_diag;
    double          scene_diag;
    /* ---- motor bodies (BodyBoundaryType::Motor, body_boundary_type.h;
     * ABD gradient/Hessian setup_abd_system_gradient_and_hessian.cu:262-339,
     * kinetic energy cal_abd_energy.cu:53-104) ----
     * an affine body whose vertices all have btype 2 is driven to turn about
     * its body x axis by motor_speed * dt per step: the target q_p rotates the
     * columns of the step-start prediction A~ (A_p = A~ R_x(motor_speed dt),
     * rows a2, a3 normalised) and the energy 0.5 dq^T P dq with
     * dq = (0, 0, a2 - a2_p, a3 - a3_p) and P = motor_strength M[6:12, 6:12]
     * joins the body's kinetic term (reference Assets/scene/
     * abd_system_config.json: speed -0.02, strength 200) */
    double          motor_speed;
    double          motor_strength;
} sgipc_scene_desc;

#ifdef __cplusplus
}
#endif
#endif /* SGIPC_SCENE_H */
