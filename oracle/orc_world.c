/*
 * orc_world.c -- one Newton iteration / time step of Stiff-GIPC on the CPU.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates StiffGIPC/GIPC.cu:
 *   computeGradientAndHessian        :10012-10228 (term order, boundary zeroing)
 *   computeEnergy                    :10392-10451
 *   lineSearch                       :10573-10670
 *   solve_subIP (loop body)          :10746-10862
 *   IPC_Solver                       :10940-11131
 *   suggestKappa/upperBoundKappa/initKappa :9746-9831
 *   ground/self/CFL step bounds      :6904-6980, 7054-7217, 8821-8968
 *   _edgeTriIntersectionQuery, segTriIntersect :8188-8302, :961-1013
 *   _stepForward/_updateVelocities/_computeXTilta :7880-7920, :8016-8038
 *   cloth membrane / bending terms (orc_cloth.c) in the assembly
 *                                    :10142-10173 and energy :10411-10416
 * and the affine-body system (abd_system/): body state q, dyadic mass, q~,
 * shape energy (orc_abd.c); vertex-level contact blocks mapped to the body
 * nodes by J^T H J (write_barrier_hessian and the ABD-FEM coupling,
 * setup_abd_system_gradient_and_hessian.cu:84-124, 419-601), applied here to
 * both orientations of every block so the reduced matrix is J^T H J exactly;
 * dq mapped back to vertex displacements by J (cal_dx_from_dq); line search
 * on q with x = J q (step_forward, cal_x_from_q); q velocity update
 * (update_velocity, cal_q_tilde).
 * and StiffGIPC/ACCD.cu point_triangle_ccd / edge_edge_ccd :373-524.
 * Lagged friction (orc_friction.c) is rebuilt at every step start from the
 * current contact sets (IPC_Solver's buildFrictionSets, GIPC.cu:11016) and
 * enters the gradient, Hessian and energy like the reference's USE_FRICTION
 * path (GIPC.cu:10043-10055, 10442-10445).
 * postLineSearch's kappa doubling is not restated: in the reference its host
 * counters h_close_gpNum / h_close_cpNum are never read back from the device
 * (GIPC.cuh:69-70, GIPC.cu:8737, 9176), so the doubling never fires.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

struct orc_world
{
    sgipc_scene_desc s;
    int              nv, nt;
    double *x, *xr, *xprev, *v, *xt, *mass, *dminv, *vol, *mu, *lam, *dx, *xsave;
    int32_t *btype, *body, *part;
    uint32_t *tets, *faces, *edges, *surf;
    /* cloth: triangles (Dm^-1 2x2, area) and bending hinges */
    int       ntri, nh;
    uint32_t *tris, *hinges;
    double *  tri_dminv, *tri_area, *h_theta0, *h_len0;
    double    c_stretch, c_shear, c_bend, c_rate;
    /* affine bodies (orc_abd.c): system nodes are [FEM vertices | 4 per body]
     * (p, a1, a2, a3); ABD vertices are the suffix [nfem, nv) */
    int       nfem, nabd, nn;
    int32_t*  abd_of;                         /* per vertex: body or -1 */
    int*      bstart;                         /* body vertex ranges     */
    int*      bfix;
    int*      tet_abd;                        /* mass-only tets         */
    double *  xbar, *bM, *bq, *bqp, *bqv, *bqt, *bqs, *bG, *bV;
    int32_t*  npart;                          /* MAS part id per node   */
    double*   dy;                             /* solution, 3 nn         */
    double    dhat2, kappa, bbox_diag2, scene_diag, mean_mass;
    int       cap, npairs, nccd, ngp;
    int32_t*  pairs;
    int32_t*  ccd;
    int32_t   counts[5];
    uint32_t* gp;
    double *  gcontact, *gshape;
    /* lagged friction sets */
    int       nfr, nfg;
    int32_t*  fr_code;
    double *  fr_lam, *fr_coord, *fr_basis;
    uint32_t* fg_ids;
    double*   fg_lam;
    int       newton_k;
    double    xdelta0;
    int       in_step;
    /* animated soft targets and the animation rate (animation_fullRate,
     * animation_subRate) */
    int*      bmotor;                         /* motor bodies and targets */
    double*   bqm;
    int       nsoft;
    uint32_t* soft_idx;
    double*   soft_tgt;
    double    anim_rate, anim_sub;
};

static void begin_step(orc_world* w);
static void end_step(orc_world* w);

static void* dup(const void* p, size_t bytes)
{
    if(!p || bytes == 0)
        return NULL;
    void* q = malloc(bytes);
    memcpy(q, p, bytes);
    return q;
}

static int bt(const orc_world* w, int v) { return w->btype ? w->btype[v] : 0; }

/* ---------------------------------------------------------------- bodies */
/* q from a vertex field f by mass-weighted least squares: M^-1 sum m J^T f */
static void body_fit(const orc_world* w, int b, const double* f, double q[12])
{
    double rhs[12] = {0}, Mi[144];
    for(int v = w->bstart[b]; v < w->bstart[b + 1]; v++)
    {
        double jt[12];
        orc_abd_jt(w->xbar + 3 * v, f + 3 * v, jt);
        for(int k = 0; k < 12; k++)
            rhs[k] += w->mass[v] * jt[k];
    }
    memcpy(Mi, w->bM + 144 * b, sizeof(Mi));
    orc_inverse(12, Mi);
    for(int i = 0; i < 12; i++)
    {
        double a = 0.0;
        for(int k = 0; k < 12; k++)
            a += Mi[i * 12 + k] * rhs[k];
        q[i] = a;
    }
}

/* vertex field of body b from q: f = J q */
static void body_apply(const orc_world* w, int b, const double q[12], double* f)
{
    for(int v = w->bstart[b]; v < w->bstart[b + 1]; v++)
        orc_abd_x(w->xbar + 3 * v, q, f + 3 * v);
}

/* x, x_prev, v, x~ of the ABD vertices from the body state */
static void bodies_to_vertices(orc_world* w)
{
    for(int b = 0; b < w->nabd; b++)
    {
        body_apply(w, b, w->bq + 12 * b, w->x);
        body_apply(w, b, w->bqp + 12 * b, w->xprev);
        body_apply(w, b, w->bqt + 12 * b, w->xt);
        body_apply(w, b, w->bqv + 12 * b, w->v);
    }
}

/* body state from vertex positions x (step boundary: q_prev = q), velocities
 * v and predictions xt */
static void vertices_to_bodies(orc_world* w, const double* v, const double* xt)
{
    for(int b = 0; b < w->nabd; b++)
    {
        body_fit(w, b, w->x, w->bq + 12 * b);
        memcpy(w->bqp + 12 * b, w->bq + 12 * b, sizeof(double) * 12);
        body_fit(w, b, v, w->bqv + 12 * b);
        body_fit(w, b, xt, w->bqt + 12 * b);
    }
    bodies_to_vertices(w);
}

/* motor target of a body from its prediction q~ (the reference's
 * q_p = J^-1 J_delta q~ + q~ with J, J_delta the affine maps of the points
 * 0, e_x, e_y, e_z and of their displacement under R_x(theta),
 * setup_abd_system_gradient_and_hessian.cu:264-307, which is A_p = A~ R_x
 * on the rows a_i), rows a2 and a3 normalised; p is not driven */
void orc_abd_motor_target(const double qt[12], double theta, double qp[12])
{
    double c = cos(theta), s = sin(theta);
    for(int k = 0; k < 3; k++)
        qp[k] = 2.0 * qt[k];
    for(int i = 0; i < 3; i++)
    {
        const double* a = qt + 3 + 3 * i;
        qp[3 + 3 * i]     = a[0];
        qp[3 + 3 * i + 1] = c * a[1] + s * a[2];
        qp[3 + 3 * i + 2] = -s * a[1] + c * a[2];
    }
    for(int i = 1; i < 3; i++)
    {
        double* a = qp + 3 + 3 * i;
        double  n = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        for(int k = 0; k < 3; k++)
            a[k] /= n;
    }
}

/* motor term of body b: dq = q - q_p on a2, a3 (zero on p, a1) and
 * P = motor_strength M[6:12, 6:12]; g += P dq, returns 0.5 dq^T P dq */
static double motor_term(const orc_world* w, int b, double* g, double* H)
{
    const double* M = w->bM + 144 * b;
    double        dq[6], e = 0.0, st = w->s.motor_strength;
    for(int k = 0; k < 6; k++)
        dq[k] = w->bq[12 * b + 6 + k] - w->bqm[12 * b + 6 + k];
    for(int r = 0; r < 6; r++)
    {
        double pr = 0.0;
        for(int c = 0; c < 6; c++)
        {
            double P = st * M[(6 + r) * 12 + 6 + c];
            pr += P * dq[c];
            if(H)
                H[(6 + r) * 12 + 6 + c] += P;
        }
        if(g)
            g[6 + r] += pr;
        e += 0.5 * dq[r] * pr;
    }
    return e;
}

/* node map of a vertex: FEM -> itself with C = I; ABD -> the 4 body nodes with
 * C_0 = I, C_(1+k) = xbar e_k^T (node a_k collects f_k xbar) */
static int node_map(const orc_world* w, int v, double C[4][9])
{
    memset(C, 0, sizeof(double) * 36);
    for(int i = 0; i < 3; i++)
        C[0][4 * i] = 1.0;
    if(w->abd_of[v] < 0)
        return 1;
    const double* xb = w->xbar + 3 * v;
    for(int k = 0; k < 3; k++)
        for(int l = 0; l < 3; l++)
            C[1 + k][l * 3 + k] = xb[l];
    return 4;
}
static int node_base(const orc_world* w, int v)
{
    return w->abd_of[v] < 0 ? v : w->nfem + 4 * w->abd_of[v];
}

/* ---------------------------------------------------------------- contacts */
static void build_cp(orc_world* w)
{
    w->npairs = orc_contact_pairs_bvh(&w->s, w->x, w->dhat2, w->pairs, w->cap, w->counts);
    if(w->npairs > w->cap)
        w->npairs = w->cap;
    w->ngp = 0;
    if(w->s.use_ground)
    {
        const double* n = w->s.ground_normal;
        for(int i = 0; i < w->s.n_surf_verts; i++)
        {
            int           v = (int)w->surf[i];
            const double* p = w->x + 3 * v;
            double        d = n[0] * p[0] + n[1] * p[1] + n[2] * p[2] - w->s.ground_offset;
            if(d * d > w->dhat2)
                continue;
            w->gp[w->ngp++] = (uint32_t)v;
        }
    }
}

/* The reference stores, next to each proximity code, the raw candidate that
 * produced it (_ccd_collisionPair): (-v-1, f0, f1, f2) for a vertex-face test
 * and (e0, e1, e2, e3) for an edge-edge test.  The oracle recomputes them by
 * re-running the candidate enumeration with a flag that records candidates. */
typedef struct
{
    int32_t* out;
    int      n, cap;
} cand_t;

static void cand_push(cand_t* c, int32_t a, int32_t b, int32_t d, int32_t e)
{
    if(c->n < c->cap)
    {
        c->out[4 * c->n + 0] = a;
        c->out[4 * c->n + 1] = b;
        c->out[4 * c->n + 2] = d;
        c->out[4 * c->n + 3] = e;
    }
    c->n++;
}

static int ovl(const double* l, const double* r, double gap)
{
    for(int c = 0; c < 3; c++)
        if((r[3 + c] - l[c]) >= gap || (l[3 + c] - r[c]) >= gap)
            return 0;
    return 1;
}

static int body_of(const orc_world* w, int v) { return w->body ? w->body[v] : -1; }

/* swept-box candidates (ConstructFullCCD + _selfQuery_*_ccd, mlbvh.cu:1410-1740) */
static int full_ccd_candidates(orc_world* w, double alpha, int32_t* out, int cap)
{
    cand_t          cd  = {out, 0, cap};
    double          gap = sqrt(w->dhat2);
    const double*   X   = w->x;
    const double*   D   = w->dx;
    uint32_t        stack[64];
    const sgipc_scene_desc* s = &w->s;
    if(s->n_faces > 0)
    {
        int       n    = s->n_faces;
        double*   leaf = (double*)malloc(sizeof(double) * 6 * n);
        uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * n);
        uint32_t* nd   = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (2 * n - 1));
        double*   bv   = (double*)malloc(sizeof(double) * 6 * (2 * n - 1));
        for(int e = 0; e < n; e++)
        {
            double up[3] = {-1e32, -1e32, -1e32}, lo[3] = {1e32, 1e32, 1e32};
            for(int k = 0; k < 3; k++)
            {
                int v = (int)s->faces[3 * e + k];
                for(int c = 0; c < 3; c++)
                {
                    double a = X[3 * v + c], b = X[3 * v + c] - D[3 * v + c] * alpha;
                    lo[c]    = (a < lo[c]) ? a : lo[c];
                    up[c]    = (up[c] < a) ? a : up[c];
                    lo[c]    = (b < lo[c]) ? b : lo[c];
                    up[c]    = (up[c] < b) ? b : up[c];
                }
            }
            for(int c = 0; c < 3; c++)
            {
                leaf[6 * e + c]     = up[c];
                leaf[6 * e + 3 + c] = lo[c];
            }
        }
        orc_lbvh_build_boxes(leaf, n, keys, nd, bv);
        for(int i = 0; i < s->n_surf_verts; i++)
        {
            int    v = (int)s->surf_verts[i];
            double q[6];
            for(int c = 0; c < 3; c++)
            {
                double a = X[3 * v + c], b = X[3 * v + c] - D[3 * v + c] * alpha;
                q[c]     = a;
                q[3 + c] = a;
                q[3 + c] = (b < q[3 + c]) ? b : q[3 + c];
                q[c]     = (q[c] < b) ? b : q[c];
            }
            int sp      = 0;
            stack[sp++] = 0;
            if(n == 1)
                sp = 0;
            do
            {
                if(n == 1)
                {
                    if(ovl(q, bv, gap))
                    {
                        const uint32_t* F = s->faces;
                        if(((body_of(w, v) != body_of(w, F[0])) || body_of(w, v) == -1)
                           && !(bt(w, v) >= 2 && bt(w, F[0]) >= 2 && bt(w, F[1]) >= 2
                                && bt(w, F[2]) >= 2)
                           && v != (int)F[0] && v != (int)F[1] && v != (int)F[2])
                            cand_push(&cd, -v - 1, F[0], F[1], F[2]);
                    }
                    break;
                }
                uint32_t node  = stack[--sp];
                uint32_t ch[2] = {nd[4 * node + 1], nd[4 * node + 2]};
                for(int c = 0; c < 2; c++)
                {
                    uint32_t C = ch[c];
                    if(!ovl(q, bv + 6 * C, gap))
                        continue;
                    uint32_t obj = nd[4 * C + 3];
                    if(obj != 0xFFFFFFFFu)
                    {
                        const uint32_t* F = s->faces + 3 * obj;
                        if(((body_of(w, v) != body_of(w, F[0])) || body_of(w, v) == -1)
                           && !(bt(w, v) >= 2 && bt(w, F[0]) >= 2 && bt(w, F[1]) >= 2
                                && bt(w, F[2]) >= 2)
                           && v != (int)F[0] && v != (int)F[1] && v != (int)F[2])
                            cand_push(&cd, -v - 1, F[0], F[1], F[2]);
                    }
                    else if(sp < 64)
                        stack[sp++] = C;
                }
            } while(sp > 0);
        }
        free(leaf);
        free(keys);
        free(nd);
        free(bv);
    }
    if(s->n_edges > 1)
    {
        int       n    = s->n_edges;
        double*   leaf = (double*)malloc(sizeof(double) * 6 * n);
        uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * n);
        uint32_t* nd   = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (2 * n - 1));
        double*   bv   = (double*)malloc(sizeof(double) * 6 * (2 * n - 1));
        for(int e = 0; e < n; e++)
        {
            double up[3] = {-1e32, -1e32, -1e32}, lo[3] = {1e32, 1e32, 1e32};
            for(int k = 0; k < 2; k++)
            {
                int v = (int)s->edges[2 * e + k];
                for(int c = 0; c < 3; c++)
                {
                    double a = X[3 * v + c], b = X[3 * v + c] - D[3 * v + c] * alpha;
                    lo[c]    = (a < lo[c]) ? a : lo[c];
                    up[c]    = (up[c] < a) ? a : up[c];
                    lo[c]    = (b < lo[c]) ? b : lo[c];
                    up[c]    = (up[c] < b) ? b : up[c];
                }
            }
            for(int c = 0; c < 3; c++)
            {
                leaf[6 * e + c]     = up[c];
                leaf[6 * e + 3 + c] = lo[c];
            }
        }
        orc_lbvh_build_boxes(leaf, n, keys, nd, bv);
        for(int i = 0; i < n; i++)
        {
            int           lf   = i + n - 1;
            const double* q    = bv + 6 * lf;
            int           self = (int)nd[4 * lf + 3];
            const uint32_t* A  = s->edges + 2 * self;
            int           sp   = 0;
            stack[sp++]        = 0;
            do
            {
                uint32_t node  = stack[--sp];
                uint32_t ch[2] = {nd[4 * node + 1], nd[4 * node + 2]};
                for(int c = 0; c < 2; c++)
                {
                    uint32_t C = ch[c];
                    if(!ovl(q, bv + 6 * C, gap))
                        continue;
                    uint32_t obj = nd[4 * C + 3];
                    if(obj != 0xFFFFFFFFu)
                    {
                        const uint32_t* B = s->edges + 2 * obj;
                        if((int)obj != self
                           && ((body_of(w, A[0]) != body_of(w, B[0])) || body_of(w, A[0]) == -1)
                           && !(A[0] == B[0] || A[0] == B[1] || A[1] == B[0] || A[1] == B[1]
                                || (int)obj < self)
                           && !(bt(w, A[0]) >= 2 && bt(w, A[1]) >= 2 && bt(w, B[0]) >= 2
                                && bt(w, B[1]) >= 2))
                            cand_push(&cd, A[0], A[1], B[0], B[1]);
                    }
                    else if(sp < 64)
                        stack[sp++] = C;
                }
            } while(sp > 0);
        }
        free(leaf);
        free(keys);
        free(nd);
        free(bv);
    }
    return cd.n;
}

/* -------------------------------------------------------------------- ACCD */
typedef struct
{
    double x, y, z;
} v3;
static v3 V(const double* p)
{
    v3 r = {p[0], p[1], p[2]};
    return r;
}
static v3 vadd(v3 a, v3 b)
{
    v3 r = {a.x + b.x, a.y + b.y, a.z + b.z};
    return r;
}
static v3 vsub(v3 a, v3 b)
{
    v3 r = {a.x - b.x, a.y - b.y, a.z - b.z};
    return r;
}
static v3 vmul(v3 a, double s)
{
    v3 r = {a.x * s, a.y * s, a.z * s};
    return r;
}
static double vsq(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static double dmax(double a, double b) { return (a < b) ? b : a; }
static double dmin(double a, double b) { return (b < a) ? b : a; }

static double pt_dist_unclassified(v3 p, v3 t0, v3 t1, v3 t2)
{
    double P[3] = {p.x, p.y, p.z}, A[3] = {t0.x, t0.y, t0.z}, B[3] = {t1.x, t1.y, t1.z},
           C[3] = {t2.x, t2.y, t2.z};
    switch(orc_dtype_pt(P, A, B, C))
    {
        case 0: return orc_d_pp(P, A);
        case 1: return orc_d_pp(P, B);
        case 2: return orc_d_pp(P, C);
        case 3: return orc_d_pe(P, A, B);
        case 4: return orc_d_pe(P, B, C);
        case 5: return orc_d_pe(P, C, A);
        case 6: return orc_d_pt(P, A, B, C);
        default: return 1e32;
    }
}

static double ee_dist_unclassified(v3 a0, v3 a1, v3 b0, v3 b1)
{
    double A0[3] = {a0.x, a0.y, a0.z}, A1[3] = {a1.x, a1.y, a1.z}, B0[3] = {b0.x, b0.y, b0.z},
           B1[3] = {b1.x, b1.y, b1.z};
    switch(orc_dtype_ee(A0, A1, B0, B1))
    {
        case 0: return orc_d_pp(A0, B0);
        case 1: return orc_d_pp(A0, B1);
        case 2: return orc_d_pe(A0, B0, B1);
        case 3: return orc_d_pp(A1, B0);
        case 4: return orc_d_pp(A1, B1);
        case 5: return orc_d_pe(A1, B0, B1);
        case 6: return orc_d_pe(B0, A0, A1);
        case 7: return orc_d_pe(B1, A0, A1);
        case 8: return orc_d_ee(A0, A1, B0, B1);
        default: return 1e32;
    }
}

static double pt_ccd(v3 p, v3 t0, v3 t1, v3 t2, v3 dp, v3 dt0, v3 dt1, v3 dt2, double eta,
                     double th)
{
    v3 mov = vmul(vadd(vadd(dt0, dt1), vadd(dt2, dp)), -0.25);
    dt0 = vadd(dt0, mov);
    dt1 = vadd(dt1, mov);
    dt2 = vadd(dt2, mov);
    dp  = vadd(dp, mov);
    double m0 = vsq(dt0), m1 = vsq(dt1), m2 = vsq(dt2);
    double maxd = sqrt(vsq(dp)) + sqrt(dmax(m0, dmax(m1, m2)));
    if(maxd == 0)
        return 1.0;
    double d2  = pt_dist_unclassified(p, t0, t1, t2);
    double d   = sqrt(d2);
    double gap = eta * (d2 - th * th) / (d + th);
    double toc = 0.0;
    int    cnt = 0;
    while(1)
    {
        cnt++;
        if(cnt > 50000)
            return toc;
        double lb = (1 - eta) * (d2 - th * th) / ((d + th) * maxd);
        p         = vadd(p, vmul(dp, lb));
        t0        = vadd(t0, vmul(dt0, lb));
        t1        = vadd(t1, vmul(dt1, lb));
        t2        = vadd(t2, vmul(dt2, lb));
        d2        = pt_dist_unclassified(p, t0, t1, t2);
        d         = sqrt(d2);
        if(toc && ((d2 - th * th) / (d + th) < gap))
            break;
        toc += lb;
        if(toc > 1.0)
            return 1.0;
    }
    return toc;
}

static double ee_ccd(v3 a0, v3 a1, v3 b0, v3 b1, v3 da0, v3 da1, v3 db0, v3 db1, double eta,
                     double th)
{
    v3 mov = vmul(vadd(vadd(da0, da1), vadd(db0, db1)), -0.25);
    da0 = vadd(da0, mov);
    da1 = vadd(da1, mov);
    db0 = vadd(db0, mov);
    db1 = vadd(db1, mov);
    double maxd = sqrt(dmax(vsq(da0), vsq(da1))) + sqrt(dmax(vsq(db0), vsq(db1)));
    if(maxd == 0)
        return 1.0;
    double d2 = ee_dist_unclassified(a0, a1, b0, b1);
    double dF = d2 - th * th;
    if(dF <= 0)
    {
        d2 = dmin(dmin(vsq(vsub(a0, b0)), vsq(vsub(a0, b1))),
                  dmin(vsq(vsub(a1, b0)), vsq(vsub(a1, b1))));
        dF = d2 - th * th;
    }
    double d   = sqrt(d2);
    double gap = eta * dF / (d + th);
    double toc = 0.0;
    int    cnt = 0;
    while(1)
    {
        cnt++;
        if(cnt > 50000)
            return toc;
        double lb = (1 - eta) * dF / ((d + th) * maxd);
        a0        = vadd(a0, vmul(da0, lb));
        a1        = vadd(a1, vmul(da1, lb));
        b0        = vadd(b0, vmul(db0, lb));
        b1        = vadd(b1, vmul(db1, lb));
        d2        = ee_dist_unclassified(a0, a1, b0, b1);
        dF        = d2 - th * th;
        if(dF <= 0)
        {
            d2 = dmin(dmin(vsq(vsub(a0, b0)), vsq(vsub(a0, b1))),
                      dmin(vsq(vsub(a1, b0)), vsq(vsub(a1, b1))));
            dF = d2 - th * th;
        }
        d = sqrt(d2);
        if(toc && (dF / (d + th) < gap))
            break;
        toc += lb;
        if(toc > 1.0)
            return 1.0;
    }
    return toc;
}

static double self_bound(orc_world* w, double slackness, const int32_t* cand, int n)
{
    if(n < 1)
        return 1.0;
    double mx  = 0.0;
    double eta = 1.0 - slackness;
    int    first = 1;
    for(int i = 0; i < n; i++)
    {
        const int32_t* c = cand + 4 * i;
        double         t;
        const double*  X = w->x;
        const double*  D = w->dx;
        if(c[0] < 0)
        {
            int p = -c[0] - 1;
            t     = 1.0
                / pt_ccd(V(X + 3 * p), V(X + 3 * c[1]), V(X + 3 * c[2]), V(X + 3 * c[3]),
                         vmul(V(D + 3 * p), -1), vmul(V(D + 3 * c[1]), -1),
                         vmul(V(D + 3 * c[2]), -1), vmul(V(D + 3 * c[3]), -1), eta, 0);
        }
        else
        {
            t = 1.0
                / ee_ccd(V(X + 3 * c[0]), V(X + 3 * c[1]), V(X + 3 * c[2]), V(X + 3 * c[3]),
                         vmul(V(D + 3 * c[0]), -1), vmul(V(D + 3 * c[1]), -1),
                         vmul(V(D + 3 * c[2]), -1), vmul(V(D + 3 * c[3]), -1), eta, 0);
        }
        if(first || mx < t)
            mx = t;
        first = 0;
    }
    return 1.0 / mx;
}

static double ground_bound(orc_world* w, double slackness)
{
    int n = w->s.n_surf_verts;
    if(n < 1)
        return 1.0;
    const double* nn = w->s.ground_normal;
    double        mx = 0.0;
    for(int i = 0; i < n; i++)
    {
        int           v    = (int)w->surf[i];
        double        temp = 1.0;
        const double* d    = w->dx + 3 * v;
        double        coef = nn[0] * d[0] + nn[1] * d[1] + nn[2] * d[2];
        if(coef > 0.0)
        {
            const double* p    = w->x + 3 * v;
            double        dist = nn[0] * p[0] + nn[1] * p[1] + nn[2] * p[2] - w->s.ground_offset;
            temp               = coef / (dist * slackness);
        }
        if(i == 0 || mx < temp)
            mx = temp;
    }
    return 1.0 / mx;
}

static double cfl_speed(orc_world* w)
{
    double mx = 0.0;
    for(int i = 0; i < w->s.n_surf_verts; i++)
    {
        const double* d = w->dx + 3 * w->surf[i];
        double        s = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if(i == 0 || mx < s)
            mx = s;
    }
    return mx;
}

/* ----------------------------------------------------------- intersection */
static int seg_tri(v3 e0, v3 e1, v3 t0, v3 t1, v3 t2)
{
    v3 c0 = vsub(t1, t0), c1 = vsub(t2, t0), c2 = vsub(e0, e1);
    v3 n  = {c0.y * c1.z - c0.z * c1.y, c0.z * c1.x - c0.x * c1.z, c0.x * c1.y - c0.y * c1.x};
    v3 a = vsub(e0, t0), b2 = vsub(e1, t0);
    if((n.x * a.x + n.y * a.y + n.z * a.z) * (n.x * b2.x + n.y * b2.y + n.z * b2.z) > 0)
        return 0;
#define DETC(A, B, C)                                                                         \
    ((A).x * (B).y * (C).z + (A).y * (B).z * (C).x + (A).z * (B).x * (C).y                   \
     - (A).z * (B).y * (C).x - (A).x * (B).z * (C).y - (B).x * (A).y * (C).z)
    double det = DETC(c0, c1, c2);
    /* segTriIntersect's |det| < 1e-20, plus a scale-relative test so that an
     * edge lying in the triangle's plane (coplanar surface faces) is rejected
     * whatever the mesh size; the absolute constant alone lets round-off
     * determinants through at centimetre scales */
    double sc = sqrt((c0.x * c0.x + c0.y * c0.y + c0.z * c0.z) * (c1.x * c1.x + c1.y * c1.y + c1.z * c1.z)
                     * (c2.x * c2.x + c2.y * c2.y + c2.z * c2.z));
    if(fabs(det) < 1e-20 || fabs(det) <= 1e-12 * sc)
        return 0;
    v3     b = vsub(e0, t0);
    double u = DETC(b, c1, c2) / det, vv = DETC(c0, b, c2) / det, t = DETC(c0, c1, b) / det;
#undef DETC
    return (u >= 0.0 && vv >= 0.0 && u + vv <= 1.0 && t >= 0.0 && t <= 1.0);
}

static int edge_tri_intersect(orc_world* w)
{
    const sgipc_scene_desc* s = &w->s;
    int                     n = s->n_edges;
    if(n < 1 || s->n_faces < 1)
        return 0;
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * n);
    uint32_t* nd   = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (2 * n - 1));
    double*   bv   = (double*)malloc(sizeof(double) * 6 * (2 * n - 1));
    orc_lbvh_build(w->x, s->edges, 2, n, keys, nd, bv);
    int      hit = 0;
    uint32_t stack[64];
    for(int f = 0; f < s->n_faces && !hit; f++)
    {
        const uint32_t* F = s->faces + 3 * f;
        double          q[6];
        orc_leaf_aabbs(w->x, F, 3, 1, q);
        int sp      = 0;
        stack[sp++] = 0;
        if(n == 1)
            sp = 0;
        while(sp > 0 && !hit)
        {
            uint32_t node  = stack[--sp];
            uint32_t ch[2] = {nd[4 * node + 1], nd[4 * node + 2]};
            for(int c = 0; c < 2 && !hit; c++)
            {
                uint32_t C = ch[c];
                if(!ovl(q, bv + 6 * C, 0.0))
                    continue;
                uint32_t obj = nd[4 * C + 3];
                if(obj != 0xFFFFFFFFu)
                {
                    const uint32_t* E = s->edges + 2 * obj;
                    if(F[0] == E[0] || F[0] == E[1] || F[1] == E[0] || F[1] == E[1]
                       || F[2] == E[0] || F[2] == E[1])
                        continue;
                    if(bt(w, F[0]) >= 2 && bt(w, F[1]) >= 2 && bt(w, F[2]) >= 2
                       && bt(w, E[0]) >= 2 && bt(w, E[1]) >= 2)
                        continue;
                    if(seg_tri(V(w->x + 3 * E[0]), V(w->x + 3 * E[1]), V(w->x + 3 * F[0]),
                               V(w->x + 3 * F[1]), V(w->x + 3 * F[2])))
                        hit = 1;
                }
                else if(sp < 64)
                    stack[sp++] = C;
            }
        }
    }
    free(keys);
    free(nd);
    free(bv);
    return hit;
}

static int is_intersected(orc_world* w)
{
    if(w->s.use_ground)
    {
        const double* n = w->s.ground_normal;
        for(int i = 0; i < w->ngp; i++)
        {
            const double* p = w->x + 3 * w->gp[i];
            if(n[0] * p[0] + n[1] * p[1] + n[2] * p[2] - w->s.ground_offset < 0)
                return 1;
        }
    }
    return edge_tri_intersect(w);
}

/* ------------------------------------------------------------------ world */
orc_world* orc_world_create(const sgipc_scene_desc* s)
{
    orc_world* w = (orc_world*)calloc(1, sizeof(orc_world));
    w->s         = *s;
    int nv = s->n_verts, nt = s->n_tets;
    w->nv = nv;
    w->nt = nt;
    size_t b3 = sizeof(double) * 3 * nv;
    w->x      = (double*)dup(s->x, b3);
    w->xr     = (double*)dup(s->x_rest ? s->x_rest : s->x, b3);
    w->xprev  = (double*)dup(s->x, b3);
    w->v      = (double*)calloc(3 * nv, sizeof(double));
    if(s->v)
        memcpy(w->v, s->v, b3);
    w->xt     = (double*)calloc(3 * nv, sizeof(double));
    w->dx     = (double*)calloc(3 * nv, sizeof(double));
    w->xsave  = (double*)calloc(3 * nv, sizeof(double));
    w->mass   = (double*)calloc(nv, sizeof(double));
    w->btype  = (int32_t*)dup(s->btype, sizeof(int32_t) * nv);
    w->body   = (int32_t*)dup(s->body_id, sizeof(int32_t) * nv);
    w->part   = (int32_t*)dup(s->part_id, sizeof(int32_t) * nv);
    w->tets   = (uint32_t*)dup(s->tets, sizeof(uint32_t) * 4 * nt);
    w->faces  = (uint32_t*)dup(s->faces, sizeof(uint32_t) * 3 * s->n_faces);
    w->edges  = (uint32_t*)dup(s->edges, sizeof(uint32_t) * 2 * s->n_edges);
    w->surf   = (uint32_t*)dup(s->surf_verts, sizeof(uint32_t) * s->n_surf_verts);
    w->s.x = w->x;
    w->s.x_rest = w->xr;
    w->s.btype = w->btype;
    w->s.body_id = w->body;
    w->s.part_id = w->part;
    w->s.tets = w->tets;
    w->s.faces = w->faces;
    w->s.edges = w->edges;
    w->s.surf_verts = w->surf;
    w->nsoft    = s->n_soft > 0 ? s->n_soft : 0;
    w->soft_idx = (uint32_t*)dup(s->soft_index, sizeof(uint32_t) * w->nsoft);
    w->soft_tgt = (double*)dup(s->soft_target, sizeof(double) * 3 * w->nsoft);
    w->s.soft_index  = w->soft_idx;
    w->s.soft_target = w->soft_tgt;
    w->anim_sub  = s->motion_rate > 0.0 ? 1.0 / s->motion_rate : 1.0;
    w->anim_rate = 0.0; /* GIPC.cuh:114 */
    w->dminv  = (double*)malloc(sizeof(double) * 9 * (nt ? nt : 1));
    w->vol    = (double*)malloc(sizeof(double) * (nt ? nt : 1));
    w->mu     = (double*)malloc(sizeof(double) * (nt ? nt : 1));
    w->lam    = (double*)malloc(sizeof(double) * (nt ? nt : 1));
    orc_fem_precompute(nv, w->xr, nt, w->tets, s->tet_E, s->tet_nu, s->tet_rho, w->dminv,
                       w->vol, w->mu, w->lam, w->mass);
    /* cloth (orc_cloth.c header: area and lumped-mass conventions) */
    w->ntri      = s->n_tris;
    w->tris      = (uint32_t*)dup(s->tris, sizeof(uint32_t) * 3 * w->ntri);
    w->s.tris    = w->tris;
    w->tri_dminv = (double*)malloc(sizeof(double) * 4 * (w->ntri ? w->ntri : 1));
    w->tri_area  = (double*)malloc(sizeof(double) * (w->ntri ? w->ntri : 1));
    orc_cloth_params(s->cloth_E, s->cloth_nu, s->cloth_bend_E, s->cloth_thickness,
                     s->cloth_strain_rate, &w->c_stretch, &w->c_shear, &w->c_bend);
    w->c_rate = s->cloth_strain_rate;
    for(int f = 0; f < w->ntri; f++)
    {
        orc_tri_precompute(w->xr, w->tris + 3 * f, w->tri_dminv + 4 * f, w->tri_area + f);
        double m = w->tri_area[f] * s->cloth_thickness * s->cloth_density / 3.0;
        for(int k = 0; k < 3; k++)
            w->mass[w->tris[3 * f + k]] += m;
    }
    w->nh       = orc_bend_hinges(w->ntri, w->tris, NULL);
    w->hinges   = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (w->nh ? w->nh : 1));
    w->h_theta0 = (double*)malloc(sizeof(double) * (w->nh ? w->nh : 1));
    w->h_len0   = (double*)malloc(sizeof(double) * (w->nh ? w->nh : 1));
    orc_bend_hinges(w->ntri, w->tris, w->hinges);
    for(int h = 0; h < w->nh; h++)
    {
        const uint32_t* q = w->hinges + 4 * h;
        w->h_theta0[h]    = orc_dihedral(w->xr + 3 * q[0], w->xr + 3 * q[1], w->xr + 3 * q[2],
                                         w->xr + 3 * q[3], NULL, NULL);
        double d2         = 0.0;
        for(int c = 0; c < 3; c++)
            d2 += (w->xr[3 * q[1] + c] - w->xr[3 * q[0] + c]) * (w->xr[3 * q[1] + c] - w->xr[3 * q[0] + c]);
        w->h_len0[h] = sqrt(d2);
    }
    /* affine bodies: mass centre, body-frame coordinates, dyadic mass,
     * volume, q = (c, I), q' fitted to v, affine gravity G = M^-1 sum m J^T g
     * (abd_system.cu _setup_abd_state, _setup_J, _setup_abd_dyadic_mass,
     * _setup_abd_gravity; cal_q_tilde) */
    w->abd_of  = (int32_t*)malloc(sizeof(int32_t) * nv);
    w->nabd    = (s->abd_body && s->n_abd > 0) ? s->n_abd : 0;
    w->nfem    = nv;
    for(int v = 0; v < nv; v++)
    {
        w->abd_of[v] = w->nabd ? s->abd_body[v] : -1;
        if(w->abd_of[v] >= 0 && w->nfem == nv)
            w->nfem = v;
    }
    w->nn      = w->nfem + 4 * w->nabd;
    w->tet_abd = (int*)calloc(nt ? nt : 1, sizeof(int));
    for(int e = 0; e < nt; e++)
        w->tet_abd[e] = w->abd_of[w->tets[4 * e]] >= 0;
    int nb1    = w->nabd ? w->nabd : 1;
    w->bstart  = (int*)calloc(nb1 + 1, sizeof(int));
    w->bfix    = (int*)calloc(nb1, sizeof(int));
    w->bmotor  = (int*)calloc(nb1, sizeof(int));
    w->bqm     = (double*)calloc(12 * nb1, sizeof(double));
    w->xbar    = (double*)calloc(3 * nv, sizeof(double));
    w->bM      = (double*)calloc(144 * nb1, sizeof(double));
    w->bq      = (double*)calloc(12 * nb1, sizeof(double));
    w->bqp     = (double*)calloc(12 * nb1, sizeof(double));
    w->bqv     = (double*)calloc(12 * nb1, sizeof(double));
    w->bqt     = (double*)calloc(12 * nb1, sizeof(double));
    w->bqs     = (double*)calloc(12 * nb1, sizeof(double));
    w->bG      = (double*)calloc(12 * nb1, sizeof(double));
    w->bV      = (double*)calloc(nb1, sizeof(double));
    for(int v = w->nfem; v < nv; v++)
        w->bstart[w->abd_of[v] + 1] = v + 1;
    w->bstart[0] = w->nfem;
    for(int e = 0; e < nt; e++)
        if(w->tet_abd[e])
            w->bV[w->abd_of[w->tets[4 * e]]] += w->vol[e];
    for(int b = 0; b < w->nabd; b++)
    {
        double ms = 0.0, c[3] = {0, 0, 0};
        for(int v = w->bstart[b]; v < w->bstart[b + 1]; v++)
        {
            ms += w->mass[v];
            for(int k = 0; k < 3; k++)
                c[k] += w->mass[v] * w->x[3 * v + k];
        }
        for(int k = 0; k < 3; k++)
            c[k] /= ms;
        /* motor bodies (btype 2 on every vertex) move freely otherwise */
        w->bmotor[b] = bt(w, w->bstart[b]) == 2;
        if(w->bmotor[b] && w->btype)
            for(int v = w->bstart[b]; v < w->bstart[b + 1]; v++)
                w->btype[v] = 0;
        w->bfix[b] = bt(w, w->bstart[b]) != 0;
        double* M  = w->bM + 144 * b;
        for(int v = w->bstart[b]; v < w->bstart[b + 1]; v++)
        {
            double J[36];
            for(int k = 0; k < 3; k++)
                w->xbar[3 * v + k] = w->x[3 * v + k] - c[k];
            orc_abd_jmat(w->xbar + 3 * v, J);
            for(int i = 0; i < 12; i++)
                for(int j = 0; j < 12; j++)
                    M[i * 12 + j] += w->mass[v]
                                     * (J[i] * J[j] + J[12 + i] * J[12 + j] + J[24 + i] * J[24 + j]);
        }
        double* q = w->bq + 12 * b;
        memset(q, 0, sizeof(double) * 12);
        for(int k = 0; k < 3; k++)
        {
            q[k]         = c[k];
            q[3 + 4 * k] = 1.0;
        }
        memcpy(w->bqp + 12 * b, q, sizeof(double) * 12);
        if(!w->bfix[b])
        {
            double* gv = (double*)calloc(3 * nv, sizeof(double));
            for(int v = w->bstart[b]; v < w->bstart[b + 1]; v++)
                for(int k = 0; k < 3; k++)
                    gv[3 * v + k] = s->gravity[k];
            body_fit(w, b, gv, w->bG + 12 * b);
            free(gv);
            body_fit(w, b, w->v, w->bqv + 12 * b);
        }
        for(int k = 0; k < 12; k++)
            w->bqt[12 * b + k] = q[k] + (w->bqv[12 * b + k] * s->dt
                                         + w->bG[12 * b + k] * (s->dt * s->dt));
    }
    /* MAS parts per node: FEM vertices keep theirs, each body is one part */
    w->npart  = (int32_t*)malloc(sizeof(int32_t) * (w->nn ? w->nn : 1));
    int maxp  = -1;
    for(int v = 0; v < w->nfem; v++)
    {
        w->npart[v] = w->part ? w->part[v] : v / 16;
        maxp        = w->npart[v] > maxp ? w->npart[v] : maxp;
    }
    for(int b = 0; b < w->nabd; b++)
        for(int k = 0; k < 4; k++)
            w->npart[w->nfem + 4 * b + k] = maxp + 1 + b;
    w->dy = (double*)calloc(3 * (w->nn ? w->nn : 1), sizeof(double));

    double msum = 0.0;
    for(int v = 0; v < nv; v++)
        msum += w->mass[v];
    w->mean_mass = msum / nv;
    w->dhat2     = s->dhat * s->dhat;
    /* bboxDiagSize2 from the face-BVH scene box (GIPC.cu:8536-8540) */
    double up[3] = {-1e32, -1e32, -1e32}, lo[3] = {1e32, 1e32, 1e32};
    for(int f = 0; f < s->n_faces; f++)
        for(int k = 0; k < 3; k++)
            for(int c = 0; c < 3; c++)
            {
                double p = w->x[3 * w->faces[3 * f + k] + c];
                lo[c]    = p < lo[c] ? p : lo[c];
                up[c]    = up[c] < p ? p : up[c];
            }
    w->bbox_diag2 = (up[0] - lo[0]) * (up[0] - lo[0]) + (up[1] - lo[1]) * (up[1] - lo[1])
                    + (up[2] - lo[2]) * (up[2] - lo[2]);
    /* scene_diag over non-fixed vertices (main.cu:853-863) */
    double u2[3] = {-1e32, -1e32, -1e32}, l2[3] = {1e32, 1e32, 1e32};
    int    any   = 0;
    for(int v = 0; v < nv; v++)
        if(bt(w, v) == 0)
        {
            any = 1;
            for(int c = 0; c < 3; c++)
            {
                double p = w->x[3 * v + c];
                l2[c]    = p < l2[c] ? p : l2[c];
                u2[c]    = u2[c] < p ? p : u2[c];
            }
        }
    w->scene_diag = any ? sqrt((u2[0] - l2[0]) * (u2[0] - l2[0]) + (u2[1] - l2[1]) * (u2[1] - l2[1])
                               + (u2[2] - l2[2]) * (u2[2] - l2[2]))
                        : 0.0;
    /* the scene's own extents when given (main.cu:853-863, GIPC::init) */
    if(s->bbox_diag > 0.0)
        w->bbox_diag2 = s->bbox_diag * s->bbox_diag;
    if(s->scene_diag > 0.0)
        w->scene_diag = s->scene_diag;
    w->kappa = s->kappa > 0 ? s->kappa : 0.0;
    w->cap   = 2 * (s->n_surf_verts + s->n_edges) + 200000;
    w->pairs = (int32_t*)malloc(sizeof(int32_t) * 4 * w->cap);
    w->ccd   = (int32_t*)malloc(sizeof(int32_t) * 4 * w->cap);
    w->gp    = (uint32_t*)malloc(sizeof(uint32_t) * (s->n_surf_verts + 1));
    w->gcontact = (double*)calloc(3 * nv, sizeof(double));
    w->gshape   = (double*)calloc(3 * nv, sizeof(double));
    w->fr_code  = (int32_t*)malloc(sizeof(int32_t) * 4 * w->cap);
    w->fr_lam   = (double*)malloc(sizeof(double) * w->cap);
    w->fr_coord = (double*)malloc(sizeof(double) * 2 * w->cap);
    w->fr_basis = (double*)malloc(sizeof(double) * 6 * w->cap);
    w->fg_ids   = (uint32_t*)malloc(sizeof(uint32_t) * (s->n_surf_verts + 1));
    w->fg_lam   = (double*)malloc(sizeof(double) * (s->n_surf_verts + 1));
    /* x_tilde for the first step (main.cu:1037-1045 computeXTilta) */
    for(int v = 0; v < nv; v++)
        for(int c = 0; c < 3; c++)
            w->xt[3 * v + c] = w->xprev[3 * v + c]
                               + (w->v[3 * v + c] * s->dt
                                  + (bt(w, v) == 0 ? s->gravity[c] * (s->dt * s->dt) : 0.0));
    bodies_to_vertices(w);
    build_cp(w);
    return w;
}

void orc_world_free(orc_world* w)
{
    if(!w)
        return;
    free(w->x); free(w->xr); free(w->xprev); free(w->v); free(w->xt); free(w->dx);
    free(w->xsave); free(w->mass); free(w->btype); free(w->body); free(w->part);
    free(w->tets); free(w->faces); free(w->edges); free(w->surf); free(w->dminv);
    free(w->vol); free(w->mu); free(w->lam); free(w->pairs); free(w->ccd); free(w->gp);
    free(w->gcontact); free(w->gshape);
    free(w->fr_code); free(w->fr_lam); free(w->fr_coord); free(w->fr_basis);
    free(w->fg_ids); free(w->fg_lam);
    free(w->tris); free(w->tri_dminv); free(w->tri_area); free(w->hinges);
    free(w->h_theta0); free(w->h_len0);
    free(w->abd_of); free(w->tet_abd); free(w->bstart); free(w->bfix); free(w->xbar);
    free(w->bM); free(w->bq); free(w->bqp); free(w->bqv); free(w->bqt); free(w->bqs);
    free(w->bG); free(w->bV); free(w->npart); free(w->dy);
    free(w->soft_idx); free(w->soft_tgt); free(w->bmotor); free(w->bqm);
    free(w);
}

void orc_world_set_positions(orc_world* w, const double* x)
{
    memcpy(w->x, x, sizeof(double) * 3 * w->nv);
    for(int b = 0; b < w->nabd; b++)   /* ABD vertices stay affine: x = J q */
    {
        body_fit(w, b, w->x, w->bq + 12 * b);
        body_apply(w, b, w->bq + 12 * b, w->x);
    }
    build_cp(w);
}
/* vertex positions exactly as given, affine bodies refitted to them but their
 * vertices left unprojected: for comparing position-level quantities (contact
 * sets, step bounds) with another implementation at identical positions */
void orc_world_set_positions_exact(orc_world* w, const double* x)
{
    memcpy(w->x, x, sizeof(double) * 3 * w->nv);
    for(int b = 0; b < w->nabd; b++)
        body_fit(w, b, w->x, w->bq + 12 * b);
    build_cp(w);
}
int orc_world_nodes(const orc_world* w) { return w->nn; }
void orc_world_get_positions(const orc_world* w, double* x)
{
    memcpy(x, w->x, sizeof(double) * 3 * w->nv);
}
void orc_world_set_xtilde(orc_world* w, const double* xt)
{
    memcpy(w->xt, xt, sizeof(double) * 3 * w->nv);
}
void orc_world_get_xtilde(const orc_world* w, double* xt)
{
    memcpy(xt, w->xt, sizeof(double) * 3 * w->nv);
}
double orc_world_get_kappa(const orc_world* w) { return w->kappa; }
void   orc_world_set_kappa(orc_world* w, double k) { w->kappa = k; }

/* ---------------------------------------------------------------- assembly */
typedef struct
{
    int*    r;
    int*    c;
    double* v;
    int     n, cap;
} trip_t;

static void trip_push(trip_t* t, int r, int c, const double* blk)
{
    if(t->n == t->cap)
    {
        t->cap = t->cap ? 2 * t->cap : 1024;
        t->r   = (int*)realloc(t->r, sizeof(int) * t->cap);
        t->c   = (int*)realloc(t->c, sizeof(int) * t->cap);
        t->v   = (double*)realloc(t->v, sizeof(double) * 9 * t->cap);
    }
    t->r[t->n] = r;
    t->c[t->n] = c;
    memcpy(t->v + 9 * t->n, blk, sizeof(double) * 9);
    t->n++;
}

/* upper blocks of an (3m)x(3m) element matrix, transposed when the global
 * row index exceeds the column (write_triplet, GIPC.cu:83-139) */
static void push_element(trip_t* t, int m, const int* ids, const double* H, int zero_if_bnd,
                         const orc_world* w)
{
    int n = 3 * m;
    for(int a = 0; a < m; a++)
        for(int b = a; b < m; b++)
        {
            int    r = ids[a], c = ids[b];
            double blk[9];
            if(r <= c)
            {
                for(int i = 0; i < 3; i++)
                    for(int j = 0; j < 3; j++)
                        blk[i * 3 + j] = H[(3 * a + i) * n + 3 * b + j];
            }
            else
            {
                int tmp = r;
                r       = c;
                c       = tmp;
                for(int i = 0; i < 3; i++)
                    for(int j = 0; j < 3; j++)
                        blk[i * 3 + j] = H[(3 * b + i) * n + 3 * a + j];
            }
            if(zero_if_bnd && (bt(w, r) != 0 || bt(w, c) != 0))
                memset(blk, 0, sizeof(blk));
            trip_push(t, r, c, blk);
        }
}

/* eps_v^2 = fDhat dt^2, fDhat = 1e-4 bboxDiag^2 (GIPC.cu:8544, :9128) */
static double friction_eps2(const orc_world* w)
{
    return 1e-4 * w->bbox_diag2 * w->s.dt * w->s.dt;
}

/* elements per parallel chunk of the assembly and energy loops */
enum { kChunk = 16384 };

static void compute_grad_hess(orc_world* w, trip_t* T)
{
    int nv = w->nv;
    memset(w->gcontact, 0, sizeof(double) * 3 * nv);
    /* kinetic on FEM vertices; affine bodies carry M (q - q~) on their nodes */
    for(int v = 0; v < nv; v++)
        for(int c = 0; c < 3; c++)
            w->gshape[3 * v + c] =
                v < w->nfem ? (w->x[3 * v + c] - w->xt[3 * v + c]) * w->mass[v] : 0.0;
    /* soft targets (_computeSoftConstraintGradientAndHessian, GIPC.cu:6365-6409):
     * gradient d rate rate (x - t), Hessian rate rate d I (d = motion stiffness) */
    for(int i = 0; i < w->nsoft; i++)
    {
        int    v  = (int)w->soft_idx[i];
        double d  = w->s.soft_stiffness, r = w->anim_rate;
        double gk = d * r * r, hk = r * r * d;
        for(int c = 0; c < 3; c++)
            w->gshape[3 * v + c] += gk * (w->x[3 * v + c] - w->soft_tgt[3 * i + c]);
        if(T)
        {
            double H[9] = {hk, 0, 0, 0, hk, 0, 0, 0, hk};
            int    ids[1] = {v};
            push_element(T, 1, ids, H, 1, w);
        }
    }
    /* barrier */
    for(int i = 0; i < w->npairs; i++)
    {
        int32_t ids[4];
        double  g[12], H[144];
        int     m = orc_barrier_grad_hess(w->x, w->xr, w->pairs + 4 * i, w->dhat2, w->kappa, ids,
                                          g, H);
        for(int a = 0; a < m; a++)
            for(int c = 0; c < 3; c++)
                w->gcontact[3 * ids[a] + c] += g[3 * a + c];
        if(T)
            push_element(T, m, ids, H, 1, w);
    }
    /* lagged friction (GIPC.cu:10043-10055: calFrictionGradient, calFrictionHessian) */
    double eps2 = friction_eps2(w);
    for(int i = 0; i < w->nfr; i++)
    {
        int32_t ids[4];
        double  g[12], H[144];
        int     m = orc_friction_grad_hess(w->x, w->xprev, w->fr_code + 4 * i, w->fr_coord + 2 * i,
                                           w->fr_basis + 6 * i, w->fr_lam[i], eps2, w->s.friction_mu,
                                           ids, g, H);
        for(int a = 0; a < m; a++)
            for(int c = 0; c < 3; c++)
                w->gcontact[3 * ids[a] + c] += g[3 * a + c];
        if(T)
            push_element(T, m, ids, H, 1, w);
    }
    for(int i = 0; i < w->nfg; i++)
    {
        int    v = (int)w->fg_ids[i];
        double g[3], H[9];
        orc_friction_ground_grad_hess(w->x + 3 * v, w->xprev + 3 * v, w->s.ground_normal,
                                      w->fg_lam[i], eps2, w->s.ground_friction_mu, g, H);
        for(int c = 0; c < 3; c++)
            w->gcontact[3 * v + c] += g[c];
        if(T)
        {
            int ids[1] = {v};
            push_element(T, 1, ids, H, 1, w);
        }
    }
    /* ground */
    for(int i = 0; i < w->ngp; i++)
    {
        int    v = (int)w->gp[i];
        double g[3], H[9];
        orc_ground_grad_hess(w->x + 3 * v, w->s.ground_normal, w->s.ground_offset, w->dhat2,
                             w->kappa, g, H);
        for(int c = 0; c < 3; c++)
            w->gcontact[3 * v + c] += g[c];
        if(T)
        {
            int ids[1] = {v};
            push_element(T, 1, ids, H, 1, w);
        }
    }
    /* FEM, cloth membrane (6 blocks per triangle) and bending (10 per hinge):
     * element gradients and Hessians of a chunk are computed in parallel,
     * then accumulated and pushed in element order */
    double* G  = (double*)malloc(sizeof(double) * 12 * kChunk);
    double* HH = (double*)malloc(sizeof(double) * 144 * kChunk);
    for(int e0 = 0; e0 < w->nt; e0 += kChunk)
    {
        int e1 = e0 + kChunk < w->nt ? e0 + kChunk : w->nt;
#pragma omp parallel for schedule(static)
        for(int e = e0; e < e1; e++)
            if(!w->tet_abd[e])
                orc_tet_grad_hess(w->x, w->tets + 4 * e, w->dminv + 9 * e, w->vol[e], w->mu[e],
                                  w->lam[e], w->s.dt, G + 12 * (e - e0), HH + 144 * (e - e0));
        for(int e = e0; e < e1; e++)
        {
            if(w->tet_abd[e])
                continue;
            const uint32_t* t = w->tets + 4 * e;
            const double*   g = G + 12 * (e - e0);
            for(int a = 0; a < 4; a++)
                for(int c = 0; c < 3; c++)
                    w->gshape[3 * t[a] + c] += g[3 * a + c];
            if(T)
            {
                int ids[4] = {(int)t[0], (int)t[1], (int)t[2], (int)t[3]};
                push_element(T, 4, ids, HH + 144 * (e - e0), 1, w);
            }
        }
    }
    for(int f0 = 0; f0 < w->ntri; f0 += kChunk)
    {
        int f1 = f0 + kChunk < w->ntri ? f0 + kChunk : w->ntri;
#pragma omp parallel for schedule(static)
        for(int f = f0; f < f1; f++)
            orc_tri_grad_hess(w->x, w->tris + 3 * f, w->tri_dminv + 4 * f, w->tri_area[f],
                              w->c_stretch, w->c_shear, w->c_rate, w->s.dt, G + 9 * (f - f0),
                              HH + 81 * (f - f0));
        for(int f = f0; f < f1; f++)
        {
            const uint32_t* t = w->tris + 3 * f;
            const double*   g = G + 9 * (f - f0);
            for(int a = 0; a < 3; a++)
                for(int c = 0; c < 3; c++)
                    w->gshape[3 * t[a] + c] += g[3 * a + c];
            if(T)
            {
                int ids[3] = {(int)t[0], (int)t[1], (int)t[2]};
                push_element(T, 3, ids, HH + 81 * (f - f0), 1, w);
            }
        }
    }
    for(int h0 = 0; h0 < w->nh; h0 += kChunk)
    {
        int h1 = h0 + kChunk < w->nh ? h0 + kChunk : w->nh;
#pragma omp parallel for schedule(static)
        for(int h = h0; h < h1; h++)
            orc_bend_grad_hess(w->x, w->hinges + 4 * h, w->h_theta0[h], w->h_len0[h], w->c_bend,
                               w->s.dt, G + 12 * (h - h0), HH + 144 * (h - h0));
        for(int h = h0; h < h1; h++)
        {
            const uint32_t* q = w->hinges + 4 * h;
            const double*   g = G + 12 * (h - h0);
            for(int a = 0; a < 4; a++)
                for(int c = 0; c < 3; c++)
                    w->gshape[3 * q[a] + c] += g[3 * a + c];
            if(T)
            {
                int ids[4] = {(int)q[0], (int)q[1], (int)q[2], (int)q[3]};
                push_element(T, 4, ids, HH + 144 * (h - h0), 1, w);
            }
        }
    }
    free(G);
    free(HH);
    /* mass diagonal of the FEM vertices (never zeroed) */
    if(T)
        for(int v = 0; v < w->nfem; v++)
        {
            double blk[9] = {w->mass[v], 0, 0, 0, w->mass[v], 0, 0, 0, w->mass[v]};
            trip_push(T, v, v, blk);
        }
}

/* dst = C_a B C_b^T (3x3 row-major) */
static void cbc(const double* Ca, const double* B, const double* Cb, double* dst)
{
    double t[9];
    for(int i = 0; i < 3; i++)
        for(int j = 0; j < 3; j++)
            t[i * 3 + j] = Ca[i * 3] * B[j] + Ca[i * 3 + 1] * B[3 + j] + Ca[i * 3 + 2] * B[6 + j];
    for(int i = 0; i < 3; i++)
        for(int j = 0; j < 3; j++)
            dst[i * 3 + j] = t[i * 3] * Cb[j * 3] + t[i * 3 + 1] * Cb[j * 3 + 1] + t[i * 3 + 2] * Cb[j * 3 + 2];
}

/* the node system: vertex-level upper blocks mapped through the node maps
 * (both orientations of off-diagonal blocks, keeping node blocks with
 * row <= col), the body blocks M + PSD(kappa V dt^2 d2 shape), the node
 * gradient [g_FEM | sum J^T g_contact + M (q - q~) + shape] */
static void reduce_to_nodes(orc_world* w, const trip_t* T, double* grad, trip_t* Y)
{
    for(int v = 0; v < w->nfem; v++)
        for(int c = 0; c < 3; c++)
            grad[3 * v + c] = bt(w, v) != 0 ? 0.0 : w->gcontact[3 * v + c] + w->gshape[3 * v + c];
    double dt2 = w->s.dt * w->s.dt;
    for(int b = 0; b < w->nabd; b++)
    {
        double* g = grad + 3 * (w->nfem + 4 * b);
        memset(g, 0, sizeof(double) * 12);
        if(w->bfix[b])
            continue;
        for(int v = w->bstart[b]; v < w->bstart[b + 1]; v++)
        {
            double jt[12];
            orc_abd_jt(w->xbar + 3 * v, w->gcontact + 3 * v, jt);
            for(int k = 0; k < 12; k++)
                g[k] += jt[k];
        }
        const double* M = w->bM + 144 * b;
        for(int i = 0; i < 12; i++)
            for(int k = 0; k < 12; k++)
                g[i] += M[i * 12 + k] * (w->bq[12 * b + k] - w->bqt[12 * b + k]);
        double sg[9], kv = w->s.abd_kappa * w->bV[b] * dt2;
        orc_abd_shape_grad(w->bq + 12 * b, sg);
        for(int k = 0; k < 9; k++)
            g[3 + k] += kv * sg[k];
        if(w->bmotor[b])
            motor_term(w, b, g, NULL);
    }
    if(!T)
        return;
    for(int t = 0; t < T->n; t++)
    {
        int    r = T->r[t], c = T->c[t];
        double Cr[4][9], Cc[4][9], blk[9], bt_[9];
        int    mr = node_map(w, r, Cr), mc = node_map(w, c, Cc);
        int    br = node_base(w, r), bc = node_base(w, c);
        for(int a = 0; a < mr; a++)
            for(int b = 0; b < mc; b++)
            {
                cbc(Cr[a], T->v + 9 * t, Cc[b], blk);
                int R = br + a, Cn = bc + b;
                if(R <= Cn)
                    trip_push(Y, R, Cn, blk);
                if(r != c && Cn <= R)
                {
                    for(int i = 0; i < 3; i++)
                        for(int j = 0; j < 3; j++)
                            bt_[i * 3 + j] = blk[j * 3 + i];
                    trip_push(Y, Cn, R, bt_);
                }
            }
    }
    for(int b = 0; b < w->nabd; b++)
    {
        double H[144];
        memcpy(H, w->bM + 144 * b, sizeof(H));
        if(!w->bfix[b])
        {
            double S[81], kv = w->s.abd_kappa * w->bV[b] * dt2;
            orc_abd_shape_hess(w->bq + 12 * b, S);
            for(int k = 0; k < 81; k++)
                S[k] *= kv;
            orc_psd_project(9, S);
            for(int i = 0; i < 9; i++)
                for(int j = 0; j < 9; j++)
                    H[(3 + i) * 12 + 3 + j] += S[i * 9 + j];
            if(w->bmotor[b])
                motor_term(w, b, NULL, H);
        }
        int base = w->nfem + 4 * b;
        for(int m = 0; m < 4; m++)
            for(int n = m; n < 4; n++)
            {
                double blk[9];
                for(int i = 0; i < 3; i++)
                    for(int j = 0; j < 3; j++)
                        blk[i * 3 + j] = H[(3 * m + i) * 12 + 3 * n + j];
                trip_push(Y, base + m, base + n, blk);
            }
    }
}

int orc_world_assemble(orc_world* w, double* grad, orc_bsr* H)
{
    trip_t T = {0}, Y = {0};
    compute_grad_hess(w, H ? &T : NULL);
    reduce_to_nodes(w, H ? &T : NULL, grad, &Y);
    int rc = 0;
    if(H)
    {
        rc       = orc_bsr_from_upper_triplets(w->nn, Y.n, Y.r, Y.c, Y.v, H);
        H->nbody = w->nabd;
    }
    free(Y.r);
    free(Y.c);
    free(Y.v);
    free(T.r);
    free(T.c);
    free(T.v);
    return rc;
}

void orc_world_energy(orc_world* w, double* terms)
{
    double kin = 0.0, ela = 0.0, bar = 0.0, gnd = 0.0, shape = 0.0, motor = 0.0;
    for(int b = 0; b < w->nabd; b++)
    {
        if(w->bfix[b])
            continue;
        const double* M = w->bM + 144 * b;
        double        d[12];
        for(int k = 0; k < 12; k++)
            d[k] = w->bq[12 * b + k] - w->bqt[12 * b + k];
        for(int i = 0; i < 12; i++)
            for(int k = 0; k < 12; k++)
                kin += 0.5 * d[i] * M[i * 12 + k] * d[k];
        if(w->bmotor[b])
            motor += motor_term(w, b, NULL, NULL);
        shape += w->s.abd_kappa * w->bV[b] * orc_abd_shape_energy(w->bq + 12 * b);
    }
    shape *= w->s.dt * w->s.dt;
    shape += motor; /* reported with the affine term (the reference counts it kinetic) */
    for(int v = 0; v < w->nfem; v++)
    {
        double d2 = 0.0;
        for(int c = 0; c < 3; c++)
        {
            double d = w->x[3 * v + c] - w->xt[3 * v + c];
            d2 += d * d;
        }
        kin += d2 * w->mass[v] * 0.5;
    }
    /* element energies in parallel, summed in element order */
    double* ebuf = (double*)malloc(sizeof(double) * kChunk);
    for(int e0 = 0; e0 < w->nt; e0 += kChunk)
    {
        int e1 = e0 + kChunk < w->nt ? e0 + kChunk : w->nt;
#pragma omp parallel for schedule(static)
        for(int e = e0; e < e1; e++)
            ebuf[e - e0] = w->tet_abd[e] ? 0.0
                                         : orc_tet_energy(w->x, w->tets + 4 * e, w->dminv + 9 * e,
                                                          w->vol[e], w->mu[e], w->lam[e]);
        for(int e = e0; e < e1; e++)
            if(!w->tet_abd[e])
                ela += ebuf[e - e0];
    }
    free(ebuf);
    ela *= w->s.dt * w->s.dt;
    for(int i = 0; i < w->npairs; i++)
        bar += orc_barrier_energy(w->x, w->xr, w->pairs + 4 * i, w->dhat2, w->kappa);
    const double* n = w->s.ground_normal;
    for(int i = 0; i < w->ngp; i++)
    {
        const double* p = w->x + 3 * w->gp[i];
        double        d = n[0] * p[0] + n[1] * p[1] + n[2] * p[2] - w->s.ground_offset;
        gnd += orc_ground_energy(d * d, w->dhat2);
    }
    gnd *= w->kappa;
    double eps2 = friction_eps2(w), fs = 0.0, fg = 0.0;
    for(int i = 0; i < w->nfr; i++)
        fs += orc_friction_energy(w->x, w->xprev, w->fr_code + 4 * i, w->fr_coord + 2 * i,
                                  w->fr_basis + 6 * i, w->fr_lam[i], eps2);
    for(int i = 0; i < w->nfg; i++)
        fg += orc_friction_ground_energy(w->x + 3 * w->fg_ids[i], w->xprev + 3 * w->fg_ids[i], n,
                                         w->fg_lam[i], eps2);
    double fric = w->s.friction_mu * fs + w->s.ground_friction_mu * fg;
    double cloth = 0.0;
    for(int f = 0; f < w->ntri; f++)
        cloth += orc_tri_energy(w->x, w->tris + 3 * f, w->tri_dminv + 4 * f, w->tri_area[f],
                                w->c_stretch, w->c_shear, w->c_rate);
    for(int h = 0; h < w->nh; h++)
        cloth += orc_bend_energy(w->x, w->hinges + 4 * h, w->h_theta0[h], w->h_len0[h], w->c_bend);
    cloth *= w->s.dt * w->s.dt;
    /* _computeSoftConstraintEnergy_Reduction (GIPC.cu:7527-7560): d |rate (x - t)|^2 / 2 */
    double soft = 0.0;
    for(int i = 0; i < w->nsoft; i++)
    {
        int    v   = (int)w->soft_idx[i];
        double dis = 0.0;
        for(int c = 0; c < 3; c++)
        {
            double e = (w->x[3 * v + c] - w->soft_tgt[3 * i + c]) * w->anim_rate;
            dis += e * e;
        }
        soft += w->s.soft_stiffness * dis * 0.5;
    }
    terms[0] = kin + ela + bar + gnd + fric + cloth + shape + soft;
    terms[8] = soft;
    terms[1] = kin;
    terms[2] = ela;
    terms[3] = bar;
    terms[4] = gnd;
    terms[5] = fric;
    terms[6] = cloth;
    terms[7] = shape;
}

static double energy(orc_world* w)
{
    double t[9];
    orc_world_energy(w, t);
    return t[0];
}

/* ------------------------------------------------------------------- kappa */
static double compute_H_b(double d, double dHat)
{
    double t = d - dHat;
    return (log(d / dHat) * -2.0 - t * 4.0 / d) + 1.0 / (d * d) * (t * t);
}

static void suggest_kappa(orc_world* w, double* kappa)
{
    double H_b = compute_H_b(1.0e-16 * w->bbox_diag2, w->dhat2);
    if(w->mean_mass == 0.0)
        *kappa = 1e11 / (4.0e-16 * w->bbox_diag2 * H_b);
    else
        *kappa = 1e11 * w->mean_mass / (4.0e-16 * w->bbox_diag2 * H_b);
}

static void upper_bound_kappa(orc_world* w, double* kappa)
{
    double H_b  = compute_H_b(1.0e-16 * w->bbox_diag2, w->dhat2);
    double kmax = 100 * 1e11 * w->mean_mass / (4.0e-16 * w->bbox_diag2 * H_b);
    if(w->mean_mass == 0.0)
        kmax = 100 * 1e11 / (4.0e-16 * w->bbox_diag2 * H_b);
    if(*kappa > kmax)
        *kappa = kmax;
}

static void init_kappa(orc_world* w)
{
    if(w->npairs <= 0)
        return;
    int     nv = w->nv;
    double* GE = (double*)calloc(3 * nv, sizeof(double));
    double* gc = (double*)calloc(3 * nv, sizeof(double));
    for(int v = 0; v < nv; v++)
        for(int c = 0; c < 3; c++)
            GE[3 * v + c] = (w->x[3 * v + c] - w->xt[3 * v + c]) * w->mass[v];
    for(int e = 0; e < w->nt; e++)
    {
        if(w->tet_abd[e])
            continue;
        const uint32_t* t = w->tets + 4 * e;
        double          g[12], H[144];
        orc_tet_grad_hess(w->x, t, w->dminv + 9 * e, w->vol[e], w->mu[e], w->lam[e], w->s.dt, g,
                          H);
        for(int a = 0; a < 4; a++)
            for(int c = 0; c < 3; c++)
                GE[3 * t[a] + c] += g[3 * a + c];
    }
    /* computeSoftConstraintGradient at the current animation rate */
    for(int i = 0; i < w->nsoft; i++)
    {
        int    v  = (int)w->soft_idx[i];
        double gk = w->s.soft_stiffness * w->anim_rate * w->anim_rate;
        for(int c = 0; c < 3; c++)
            GE[3 * v + c] += gk * (w->x[3 * v + c] - w->soft_tgt[3 * i + c]);
    }
    for(int i = 0; i < w->ngp; i++)
    {
        int    v = (int)w->gp[i];
        double g[3], H[9];
        orc_ground_grad_hess(w->x + 3 * v, w->s.ground_normal, w->s.ground_offset, w->dhat2, 1.0,
                             g, H);
        for(int c = 0; c < 3; c++)
            gc[3 * v + c] += g[c];
    }
    for(int i = 0; i < w->npairs; i++)
    {
        int32_t ids[4];
        double  g[12], H[144];
        int m = orc_barrier_grad_hess(w->x, w->xr, w->pairs + 4 * i, w->dhat2, 1.0, ids, g, H);
        for(int a = 0; a < m; a++)
            for(int c = 0; c < 3; c++)
                gc[3 * ids[a] + c] += g[3 * a + c];
    }
    double gsum = 0.0, gsnorm = 0.0;
    for(int i = 0; i < 3 * nv; i++)
    {
        gsum += gc[i] * GE[i];
        gsnorm += gc[i] * gc[i];
    }
    double minKappa = -gsum / gsnorm;
    if(minKappa > 0.0)
        w->kappa = minKappa;
    suggest_kappa(w, &minKappa);
    if(w->kappa < minKappa)
        w->kappa = minKappa;
    upper_bound_kappa(w, &w->kappa);
    free(GE);
    free(gc);
}

/* ------------------------------------------------------------------ Newton */
static void step_forward(orc_world* w, double alpha)
{
    for(int v = 0; v < w->nfem; v++)
        if(bt(w, v) == 0)
            for(int c = 0; c < 3; c++)
                w->x[3 * v + c] = w->xsave[3 * v + c] - w->dx[3 * v + c] * alpha;
    for(int b = 0; b < w->nabd; b++)
    {
        if(w->bfix[b])
            continue;
        const double* dq = w->dy + 3 * (w->nfem + 4 * b);
        for(int k = 0; k < 12; k++)
            w->bq[12 * b + k] = w->bqs[12 * b + k] - dq[k] * alpha;
        body_apply(w, b, w->bq + 12 * b, w->x);
    }
}

static void line_search(orc_world* w, double* alpha, double cfl_alpha, int* nls)
{
    double E0  = energy(w);
    double gdx = 0.0;
    for(int i = 0; i < 3 * w->nv; i++)
        gdx += w->gcontact[i] * w->dx[i];
    double c1m = w->s.armijo_c1 * gdx;
    memcpy(w->xsave, w->x, sizeof(double) * 3 * w->nv);
    memcpy(w->bqs, w->bq, sizeof(double) * 12 * (w->nabd ? w->nabd : 1));
    step_forward(w, *alpha);
    while(is_intersected(w))
    {
        *alpha /= 2.0;
        *alpha = dmin(cfl_alpha, *alpha);
        step_forward(w, *alpha);
    }
    build_cp(w);
    double E  = energy(w);
    double LF = *alpha;
    int    k  = 0;
    for(int i = 0; i < 30; i++)
    {
        if(E <= E0 + c1m * *alpha)
            break;
        *alpha *= w->s.armijo_beta;
        k++;
        step_forward(w, *alpha);
        build_cp(w);
        E = energy(w);
    }
    if(*alpha < LF)
    {
        int recompute = 0;
        while(is_intersected(w))
        {
            *alpha *= w->s.armijo_beta;
            double lo = w->s.armijo_alpha_min, hi = cfl_alpha;
            *alpha = (*alpha < lo) ? lo : (hi < *alpha ? hi : *alpha);
            step_forward(w, *alpha);
            recompute = 1;
        }
        if(recompute)
            build_cp(w);
    }
    if(nls)
        *nls = k;
}

/* raw candidates of the current proximity codes: the reference keeps the
 * candidate that produced each code (_ccd_collisionPair); since every code
 * determines its candidate only up to the face/edge that produced it, the
 * oracle re-enumerates candidates with the distance test applied. */
static int proximity_candidates(orc_world* w, int32_t* out, int cap)
{
    /* reuse the swept query with alpha = 0 and keep only candidates that
     * produce a proximity code */
    int    n  = full_ccd_candidates(w, 0.0, out, cap);
    int    m  = 0;
    for(int i = 0; i < n && i < cap; i++)
    {
        int32_t* c  = out + 4 * i;
        double   d2 = 1e300;
        if(c[0] < 0)
        {
            int           p = -c[0] - 1;
            const double* X = w->x;
            switch(orc_dtype_pt(X + 3 * p, X + 3 * c[1], X + 3 * c[2], X + 3 * c[3]))
            {
                case 0: d2 = orc_d_pp(X + 3 * p, X + 3 * c[1]); break;
                case 1: d2 = orc_d_pp(X + 3 * p, X + 3 * c[2]); break;
                case 2: d2 = orc_d_pp(X + 3 * p, X + 3 * c[3]); break;
                case 3: d2 = orc_d_pe(X + 3 * p, X + 3 * c[1], X + 3 * c[2]); break;
                case 4: d2 = orc_d_pe(X + 3 * p, X + 3 * c[2], X + 3 * c[3]); break;
                case 5: d2 = orc_d_pe(X + 3 * p, X + 3 * c[3], X + 3 * c[1]); break;
                case 6: d2 = orc_d_pt(X + 3 * p, X + 3 * c[1], X + 3 * c[2], X + 3 * c[3]); break;
            }
        }
        else
        {
            const double* X = w->x;
            switch(orc_dtype_ee(X + 3 * c[0], X + 3 * c[1], X + 3 * c[2], X + 3 * c[3]))
            {
                case 0: d2 = orc_d_pp(X + 3 * c[0], X + 3 * c[2]); break;
                case 1: d2 = orc_d_pp(X + 3 * c[0], X + 3 * c[3]); break;
                case 2: d2 = orc_d_pe(X + 3 * c[0], X + 3 * c[2], X + 3 * c[3]); break;
                case 3: d2 = orc_d_pp(X + 3 * c[1], X + 3 * c[2]); break;
                case 4: d2 = orc_d_pp(X + 3 * c[1], X + 3 * c[3]); break;
                case 5: d2 = orc_d_pe(X + 3 * c[1], X + 3 * c[2], X + 3 * c[3]); break;
                case 6: d2 = orc_d_pe(X + 3 * c[2], X + 3 * c[0], X + 3 * c[1]); break;
                case 7: d2 = orc_d_pe(X + 3 * c[3], X + 3 * c[0], X + 3 * c[1]); break;
                case 8: d2 = orc_d_ee(X + 3 * c[0], X + 3 * c[1], X + 3 * c[2], X + 3 * c[3]); break;
            }
        }
        if(d2 < w->dhat2)
        {
            memmove(out + 4 * m, c, sizeof(int32_t) * 4);
            m++;
        }
    }
    return m;
}

/* Step-size bounds of the direction w->dx (solve_subIP, GIPC.cu:10786-10822):
 * ground bound (slackness 0.8), ACCD bound over the proximity pairs, full-CCD
 * candidates at that step, and when there are any the CFL bound
 * sqrt(dHat) / max|dx| * 0.5, re-bounded by ACCD over the candidates when the
 * CCD step exceeds twice the CFL step.  out = {ground, self, candidates,
 * max speed (0 without candidates), alpha_cfl, alpha}. */
static void step_bounds(orc_world* w, double out[6])
{
    double alpha = 1.0, ag = 1.0;
    if(w->s.use_ground)
        ag = ground_bound(w, 0.8);
    alpha         = dmin(alpha, ag);
    int32_t* cand = w->ccd;
    int      np   = proximity_candidates(w, cand, w->cap);
    double   as   = self_bound(w, 0.8, cand, np);
    alpha         = dmin(alpha, as);
    double temp_alpha = alpha, alpha_cfl = alpha, maxSpeed = 0.0;
    int    nccd = full_ccd_candidates(w, temp_alpha, cand, w->cap);
    if(nccd > w->cap)
        nccd = w->cap;
    w->nccd = nccd;
    if(nccd > 0)
    {
        maxSpeed  = cfl_speed(w);
        alpha_cfl = sqrt(w->dhat2) / maxSpeed * 0.5;
        alpha     = dmin(alpha, alpha_cfl);
        if(temp_alpha > 2 * alpha_cfl)
        {
            alpha = dmin(temp_alpha, self_bound(w, 0.8, cand, nccd));
            alpha = dmax(alpha, alpha_cfl);
        }
    }
    out[0] = ag;
    out[1] = as;
    out[2] = nccd;
    out[3] = maxSpeed;
    out[4] = alpha_cfl;
    out[5] = alpha;
}

void orc_world_ccd_bound(orc_world* w, const double* dx, double* out)
{
    memcpy(w->dx, dx, sizeof(double) * 3 * w->nv);
    step_bounds(w, out);
}

int orc_world_ccd_candidates(orc_world* w, const double* dx, double alpha, int32_t* out, int cap)
{
    memcpy(w->dx, dx, sizeof(double) * 3 * w->nv);
    int n = full_ccd_candidates(w, alpha, w->ccd, w->cap);
    if(n > w->cap)
        n = w->cap;
    if(out)
        memcpy(out, w->ccd, sizeof(int32_t) * 4 * (n < cap ? n : cap));
    return n;
}

double orc_barrier_H_b(double d, double dhat) { return compute_H_b(d, dhat); }

void orc_world_kappa_info(orc_world* w, double* out)
{
    double k = 0.0;
    suggest_kappa(w, &k);
    out[0] = k;
    k      = 1e300;
    upper_bound_kappa(w, &k);
    out[1] = k;
    out[2] = w->mean_mass;
    out[3] = w->bbox_diag2;
    out[4] = w->dhat2;
}

double orc_world_init_kappa(orc_world* w)
{
    w->kappa = 0.0;
    init_kappa(w);
    return w->kappa;
}

int orc_world_newton_iteration(orc_world* w, double* alpha_out, int* pcg_iters_out,
                               double* xdelta_out, int* converged_out)
{
    int      nv   = w->nv;
    double*  grad = (double*)malloc(sizeof(double) * 3 * (w->nn ? w->nn : 1));
    orc_bsr  H    = {0};
    orc_world_assemble(w, grad, &H);
    double xdelta = 0.0;
    for(int i = 0; i < 3 * nv; i++)
        xdelta = dmax(xdelta, fabs(w->dx[i]));
    if(xdelta_out)
        *xdelta_out = xdelta;
    int k = w->newton_k;
    if(k == 1)
        w->xdelta0 = xdelta;
    if(k > 0
       && (xdelta < w->scene_diag * w->s.abs_xdelta_tol || xdelta < w->s.rel_xdelta_tol * w->xdelta0))
    {
        if(converged_out)
            *converged_out = 1;
        orc_bsr_free(&H);
        free(grad);
        return 0;
    }
    if(converged_out)
        *converged_out = 0;
    orc_mas* mas = NULL;
    if(w->s.precond == SGIPC_PRECOND_MAS)
        mas = orc_mas_create(&H, w->npart, w->s.mas_precision == SGIPC_MAS_FP32);
    double rz0, rz;
    int    it = orc_pcg(&H, mas, w->s.precond == SGIPC_PRECOND_MAS ? 1 : 0, grad, w->dy,
                        w->s.pcg_rel_tol, w->s.pcg_abs_tol, w->s.pcg_max_iter,
                        w->s.pcg_use_preconditioned_norm, 10, &rz0, &rz);
    /* vertex displacements: FEM from the solution, ABD vertices J dq */
    memcpy(w->dx, w->dy, sizeof(double) * 3 * w->nfem);
    for(int b = 0; b < w->nabd; b++)
    {
        const double* dq = w->dy + 3 * (w->nfem + 4 * b);
        for(int v = w->bstart[b]; v < w->bstart[b + 1]; v++)
        {
            double d[3];
            orc_abd_x(w->xbar + 3 * v, dq, d);
            for(int c = 0; c < 3; c++)
                w->dx[3 * v + c] = d[c];
        }
    }
    if(pcg_iters_out)
        *pcg_iters_out = it;
    orc_mas_free(mas);
    orc_bsr_free(&H);
    free(grad);

    double b[6];
    step_bounds(w, b);
    double alpha = b[5], alpha_cfl = b[4];
    line_search(w, &alpha, alpha_cfl, NULL);
    if(w->kappa == 0.0)
        init_kappa(w);
    w->newton_k++;
    if(alpha_out)
        *alpha_out = alpha;
    return 0;
}

/* buildFrictionSets (GIPC.cu:8601-8637) on the current contact sets */
static void build_friction_sets(orc_world* w)
{
    w->nfr = 0;
    w->nfg = 0;
    if(w->s.friction_mu != 0.0)
        for(int i = 0; i < w->npairs; i++)
        {
            double lam, c[2], B[6];
            if(orc_friction_lag(w->x, w->pairs + 4 * i, w->dhat2, w->kappa, &lam, c, B) < 0)
                continue;
            int k = w->nfr++;
            memcpy(w->fr_code + 4 * k, w->pairs + 4 * i, sizeof(int32_t) * 4);
            w->fr_lam[k] = lam;
            memcpy(w->fr_coord + 2 * k, c, sizeof(c));
            memcpy(w->fr_basis + 6 * k, B, sizeof(B));
        }
    if(w->s.ground_friction_mu != 0.0 && w->s.use_ground)
        for(int i = 0; i < w->ngp; i++)
        {
            int k         = w->nfg++;
            w->fg_ids[k]  = w->gp[i];
            w->fg_lam[k]  = orc_friction_lag_ground(w->x + 3 * w->gp[i], w->s.ground_normal,
                                                   w->s.ground_offset, w->dhat2, w->kappa);
        }
}

static void begin_step(orc_world* w)
{
    upper_bound_kappa(w, &w->kappa);
    if(w->kappa < 1e-16)
        suggest_kappa(w, &w->kappa);
    init_kappa(w);
    w->anim_rate = w->anim_sub; /* animation_fullRate = animation_subRate (GIPC.cu:11018) */
    for(int b = 0; b < w->nabd; b++)
        if(w->bmotor[b])
            orc_abd_motor_target(w->bqt + 12 * b, w->s.motor_speed * w->s.dt, w->bqm + 12 * b);
    build_friction_sets(w);
    memset(w->dx, 0, sizeof(double) * 3 * w->nv);
    memset(w->dy, 0, sizeof(double) * 3 * (w->nn ? w->nn : 1));
    w->newton_k = 0;
    w->xdelta0  = 0.0;
    w->in_step  = 1;
}

static void end_step(orc_world* w)
{
    /* bodies: q' = (q - q_prev) / dt, q_prev = q, q~ = q + dt q' + dt^2 G
     * (update_velocity, cal_q_tilde) */
    for(int b = 0; b < w->nabd; b++)
    {
        double* q  = w->bq + 12 * b;
        double* qp = w->bqp + 12 * b;
        double* qv = w->bqv + 12 * b;
        double* qt = w->bqt + 12 * b;
        for(int k = 0; k < 12; k++)
        {
            qv[k] = w->bfix[b] ? 0.0 : (q[k] - qp[k]) * (1 / w->s.dt);
            qp[k] = q[k];
            qt[k] = w->bfix[b] ? q[k]
                               : q[k] + (qv[k] * w->s.dt + w->bG[12 * b + k] * (w->s.dt * w->s.dt));
        }
    }
    for(int v = 0; v < w->nfem; v++)
        for(int c = 0; c < 3; c++)
        {
            if(bt(w, v) == 0)
                w->v[3 * v + c] = (w->x[3 * v + c] - w->xprev[3 * v + c]) * (1 / w->s.dt);
            else
                w->v[3 * v + c] = 0.0;
            w->xprev[3 * v + c] = w->x[3 * v + c];
        }
    for(int v = 0; v < w->nfem; v++)
        for(int c = 0; c < 3; c++)
            w->xt[3 * v + c] = w->xprev[3 * v + c]
                               + (w->v[3 * v + c] * w->s.dt
                                  + (bt(w, v) == 0 ? w->s.gravity[c] * (w->s.dt * w->s.dt) : 0.0));
    bodies_to_vertices(w);
    w->in_step = 0;
}

/* after a Newton loop: 1 when the animation is complete (animation_fullRate >
 * 0.99) and the step ends, else the next loop at the next rate with the lagged
 * friction sets rebuilt and a fresh solve_subIP state (GIPC.cu:11038-11079) */
static int next_newton_loop(orc_world* w)
{
    if(w->anim_rate > 0.99)
        return 1;
    w->anim_rate += w->anim_sub;
    build_friction_sets(w);
    memset(w->dx, 0, sizeof(double) * 3 * w->nv);
    memset(w->dy, 0, sizeof(double) * 3 * (w->nn ? w->nn : 1));
    w->newton_k = 0;
    w->xdelta0  = 0.0;
    return 0;
}

int orc_world_time_step(orc_world* w, int* newton_iters_out, int* pcg_iters_out)
{
    begin_step(w);
    int total_pcg = 0, total_k = 0;
    int cap = w->s.newton_max_iter > 0 ? w->s.newton_max_iter : 10000;
    do
    {
        for(int k = 0; k < cap; k++)
        {
            int it = 0, conv = 0;
            orc_world_newton_iteration(w, NULL, &it, NULL, &conv);
            if(conv)
                break;
            total_pcg += it;
            total_k++;
        }
    } while(!next_newton_loop(w));
    end_step(w);
    if(newton_iters_out)
        *newton_iters_out = total_k;
    if(pcg_iters_out)
        *pcg_iters_out = total_pcg;
    return 0;
}

/* animation rate of the soft targets (animation_fullRate), for tests */
void orc_world_set_anim_rate(orc_world* w, double r) { w->anim_rate = r; }
double orc_world_get_anim_rate(const orc_world* w) { return w->anim_rate; }

/* new targets of the soft constraints (the frame update of main.cu:1117-1140) */
void orc_world_set_soft_targets(orc_world* w, const double* t)
{
    if(w->nsoft)
        memcpy(w->soft_tgt, t, sizeof(double) * 3 * w->nsoft);
}

/* exported ACCD entry points for the oracle's own tests (displacements are
 * the motion over the step, as passed to point_triangle_ccd) */
double orc_pt_ccd(const double* p, const double* t0, const double* t1, const double* t2,
                  const double* dp, const double* dt0, const double* dt1, const double* dt2,
                  double eta, double thickness)
{
    return pt_ccd(V(p), V(t0), V(t1), V(t2), V(dp), V(dt0), V(dt1), V(dt2), eta, thickness);
}

double orc_ee_ccd(const double* a0, const double* a1, const double* b0, const double* b1,
                  const double* da0, const double* da1, const double* db0, const double* db1,
                  double eta, double thickness)
{
    return ee_ccd(V(a0), V(a1), V(b0), V(b1), V(da0), V(da1), V(db0), V(db1), eta, thickness);
}

/* state at a time-step boundary (x = x_prev), as exported by the GPU run */
void orc_world_set_state(orc_world* w, const double* x, const double* v, const double* xt,
                         double kappa)
{
    memcpy(w->x, x, sizeof(double) * 3 * w->nv);
    memcpy(w->xprev, x, sizeof(double) * 3 * w->nv);
    memcpy(w->v, v, sizeof(double) * 3 * w->nv);
    memcpy(w->xt, xt, sizeof(double) * 3 * w->nv);
    vertices_to_bodies(w, v, xt);
    w->kappa    = kappa;
    w->in_step  = 0;
    build_cp(w);
}

/* exactly n Newton iterations that solve, across time-step boundaries */
int orc_world_advance(orc_world* w, int n, int* steps_out, int* pcg_out)
{
    int done = 0, steps = 0, pcg = 0;
    int cap  = w->s.newton_max_iter > 0 ? w->s.newton_max_iter : 10000;
    while(done < n)
    {
        if(!w->in_step)
        {
            begin_step(w);
            steps++;
        }
        int it = 0, conv = 0;
        orc_world_newton_iteration(w, NULL, &it, NULL, &conv);
        if(conv)
        {
            if(next_newton_loop(w))
                end_step(w);
            continue;
        }
        done++;
        pcg += it;
        if(w->newton_k >= cap && next_newton_loop(w))
            end_step(w);
    }
    if(steps_out)
        *steps_out = steps;
    if(pcg_out)
        *pcg_out = pcg;
    return done;
}
