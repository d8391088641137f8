/*
 * orc_contact.c -- proximity classification, contact-pair enumeration and the
 * GIPC barrier.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates StiffGIPC/mlbvh.cu:
 *   _d_PP/_d_PT/_d_PE/_d_EE        :231-273 (same operation order)
 *   _compute_epx_cp                :281-288
 *   _dType_PT / _dType_EE          :290-534
 *   _checkPTintersection           :537-653 (signed int4 codes, h_cpNum)
 *   _checkEEintersection           :687-1062 with smooth == false, so the
 *                                   mollified-EE code is never produced
 *   _selfQuery_vf / _selfQuery_ee  :1308-1408, :1511-1634 (exclusions)
 * and the RANK 2 barrier of StiffGIPC/GIPC.cu:
 *   energy  kappa (d2 - dhat2)^2 ln^2(d2/dhat2)            :703-959
 *   gradient b'(s) ds/dx, Hessian lambda0 ds/dx ds/dx^T with
 *   s = d / sqrt(dhat2) (the NEWF "F" whose only active column is ds/dx,
 *   GIPC.cu:3134-3376, 3890-4135, 4393-4690, 4692-4947)   .
 * ds/dx is formed from the closest points (envelope theorem) instead of the
 * reference's generated pFpx_* code (GIPC_PDerivative.cuh:1829-2405).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

typedef struct
{
    double x, y, z;
} d3;

static d3 ld(const double* p)
{
    d3 r = {p[0], p[1], p[2]};
    return r;
}
static d3 sub(d3 a, d3 b)
{
    d3 r = {a.x - b.x, a.y - b.y, a.z - b.z};
    return r;
}
static d3 cross(d3 a, d3 b)
{
    d3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
static double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static double sqn(d3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }

/* __Determiant with columns (c0, c1, c2): m[r][c] */
static double det_cols(d3 c0, d3 c1, d3 c2)
{
    double m00 = c0.x, m01 = c1.x, m02 = c2.x;
    double m10 = c0.y, m11 = c1.y, m12 = c2.y;
    double m20 = c0.z, m21 = c1.z, m22 = c2.z;
    return m00 * m11 * m22 + m10 * m21 * m02 + m20 * m01 * m12 - m20 * m11 * m02
           - m00 * m12 * m21 - m01 * m10 * m22;
}

double orc_d_pp(const double* v0, const double* v1) { return sqn(sub(ld(v0), ld(v1))); }

double orc_d_pt(const double* v0, const double* v1, const double* v2, const double* v3)
{
    d3     b   = cross(sub(ld(v2), ld(v1)), sub(ld(v3), ld(v1)));
    double aTb = dot(sub(ld(v0), ld(v1)), b);
    return aTb * aTb / sqn(b);
}

double orc_d_pe(const double* v0, const double* v1, const double* v2)
{
    return sqn(cross(sub(ld(v1), ld(v0)), sub(ld(v2), ld(v0)))) / sqn(sub(ld(v2), ld(v1)));
}

double orc_d_ee(const double* v0, const double* v1, const double* v2, const double* v3)
{
    d3     b   = cross(sub(ld(v1), ld(v0)), sub(ld(v3), ld(v2)));
    double aTb = dot(sub(ld(v2), ld(v0)), b);
    return aTb * aTb / sqn(b);
}

static double epx(const double* r0, const double* r1, const double* r2, const double* r3)
{
    return 1e-3 * sqn(sub(ld(r0), ld(r1))) * sqn(sub(ld(r2), ld(r3)));
}

int orc_dtype_pt(const double* pv0, const double* pv1, const double* pv2, const double* pv3)
{
    d3 v0 = ld(pv0), v1 = ld(pv1), v2 = ld(pv2), v3 = ld(pv3);
    d3 basis0 = sub(v2, v1), basis1 = sub(v3, v1), basis2 = sub(v0, v1);
    d3 nVec   = cross(basis0, basis1);
    basis1    = cross(basis0, nVec);
    double D  = det_cols(basis0, basis1, nVec);
    double p0x = det_cols(basis2, basis1, nVec) / D;
    double p0y = det_cols(basis0, basis2, nVec) / D;
    if(p0x > 0 && p0x < 1 && p0y >= 0)
        return 3;
    basis0 = sub(v3, v2);
    basis1 = cross(basis0, nVec);
    basis2 = sub(v0, v2);
    D          = det_cols(basis0, basis1, nVec);
    double p1x = det_cols(basis2, basis1, nVec) / D;
    double p1y = det_cols(basis0, basis2, nVec) / D;
    if(p1x > 0.0 && p1x < 1.0 && p1y >= 0.0)
        return 4;
    basis0 = sub(v1, v3);
    basis1 = cross(basis0, nVec);
    basis2 = sub(v0, v3);
    D          = det_cols(basis0, basis1, nVec);
    double p2x = det_cols(basis2, basis1, nVec) / D;
    double p2y = det_cols(basis0, basis2, nVec) / D;
    if(p2x > 0.0 && p2x < 1.0 && p2y >= 0.0)
        return 5;
    if(p0x <= 0.0 && p2x >= 1.0)
        return 0;
    if(p1x <= 0.0 && p0x >= 1.0)
        return 1;
    if(p2x <= 0.0 && p1x >= 1.0)
        return 2;
    return 6;
}

int orc_dtype_ee(const double* pv0, const double* pv1, const double* pv2, const double* pv3)
{
    d3     v0 = ld(pv0), v1 = ld(pv1), v2 = ld(pv2), v3 = ld(pv3);
    d3     u = sub(v1, v0), v = sub(v3, v2), w = sub(v0, v2);
    double a = sqn(u), b = dot(u, v), c = sqn(v), d = dot(u, w), e = dot(v, w);
    double D = a * c - b * b, tD = D, sN, tN;
    int    defaultCase = 8;
    sN                 = (b * e - c * d);
    if(sN <= 0.0)
    {
        tN          = e;
        tD          = c;
        defaultCase = 2;
    }
    else if(sN >= D)
    {
        tN          = e + b;
        tD          = c;
        defaultCase = 5;
    }
    else
    {
        tN = (a * e - b * d);
        if(tN > 0.0 && tN < tD
           && (dot(w, cross(u, v)) == 0.0 || sqn(cross(u, v)) < 1.0e-20 * a * c))
        {
            if(sN < D / 2)
            {
                tN          = e;
                tD          = c;
                defaultCase = 2;
            }
            else
            {
                tN          = e + b;
                tD          = c;
                defaultCase = 5;
            }
        }
    }
    if(tN <= 0.0)
    {
        if(-d <= 0.0)
            return 0;
        else if(-d >= a)
            return 3;
        else
            return 6;
    }
    else if(tN >= tD)
    {
        if((-d + b) <= 0.0)
            return 1;
        else if((-d + b) >= a)
            return 4;
        else
            return 7;
    }
    return defaultCase;
}

typedef struct
{
    int32_t* pairs;
    int      cap, n;
    int32_t* counts;
} sink_t;

static void emit(sink_t* s, int32_t a, int32_t b, int32_t c, int32_t d, int type_slot)
{
    if(s->n < s->cap)
    {
        s->pairs[4 * s->n + 0] = a;
        s->pairs[4 * s->n + 1] = b;
        s->pairs[4 * s->n + 2] = c;
        s->pairs[4 * s->n + 3] = d;
    }
    s->n++;
    s->counts[0]++;
    s->counts[type_slot]++;
}

static void check_pt(const double* X, int id0, int id1, int id2, int id3, double dHat, sink_t* s)
{
    const double *v0 = X + 3 * id0, *v1 = X + 3 * id1, *v2 = X + 3 * id2, *v3 = X + 3 * id3;
    int    t = orc_dtype_pt(v0, v1, v2, v3);
    double d = 100;
    switch(t)
    {
        case 0:
            d = orc_d_pp(v0, v1);
            if(d < dHat)
                emit(s, -id0 - 1, id1, -1, -1, 2);
            break;
        case 1:
            d = orc_d_pp(v0, v2);
            if(d < dHat)
                emit(s, -id0 - 1, id2, -1, -1, 2);
            break;
        case 2:
            d = orc_d_pp(v0, v3);
            if(d < dHat)
                emit(s, -id0 - 1, id3, -1, -1, 2);
            break;
        case 3:
            d = orc_d_pe(v0, v1, v2);
            if(d < dHat)
                emit(s, -id0 - 1, id1, id2, -1, 3);
            break;
        case 4:
            d = orc_d_pe(v0, v2, v3);
            if(d < dHat)
                emit(s, -id0 - 1, id2, id3, -1, 3);
            break;
        case 5:
            d = orc_d_pe(v0, v3, v1);
            if(d < dHat)
                emit(s, -id0 - 1, id3, id1, -1, 3);
            break;
        case 6:
            d = orc_d_pt(v0, v1, v2, v3);
            if(d < dHat)
                emit(s, -id0 - 1, id1, id2, id3, 4);
            break;
        default: break;
    }
}

static void check_ee(const double* X, const double* XR, int id0, int id1, int id2, int id3,
                     int obj, double dHat, sink_t* s)
{
    const double *v0 = X + 3 * id0, *v1 = X + 3 * id1, *v2 = X + 3 * id2, *v3 = X + 3 * id3;
    const double *r0 = XR + 3 * id0, *r1 = XR + 3 * id1, *r2 = XR + 3 * id2, *r3 = XR + 3 * id3;
    int    t     = orc_dtype_ee(v0, v1, v2, v3);
    double d     = 100.0;
    int    add_e = -1;
    double cr;
    /* eeSqureNCross / eps_x decide the mollifier flag stored in .w */
#define FLAG01()                                                                  \
    cr    = sqn(cross(sub(ld(v0), ld(v1)), sub(ld(v2), ld(v3))));                 \
    add_e = (cr < epx(r0, r1, r2, r3)) ? -obj - 2 : -1;
#define FLAG23()                                                                  \
    cr    = sqn(cross(sub(ld(v2), ld(v3)), sub(ld(v0), ld(v1))));                 \
    add_e = (cr < epx(r2, r3, r0, r1)) ? -obj - 2 : -1;
    switch(t)
    {
        case 0:
            d = orc_d_pp(v0, v2);
            if(d < dHat)
            {
                FLAG01();
                emit(s, -id0 - 1, id2, -1, add_e, 2);
            }
            break;
        case 1:
            d = orc_d_pp(v0, v3);
            if(d < dHat)
            {
                FLAG01();
                emit(s, -id0 - 1, id3, -1, add_e, 2);
            }
            break;
        case 2:
            d = orc_d_pe(v0, v2, v3);
            if(d < dHat)
            {
                FLAG01();
                emit(s, -id0 - 1, id2, id3, add_e, 3);
            }
            break;
        case 3:
            d = orc_d_pp(v1, v2);
            if(d < dHat)
            {
                FLAG01();
                emit(s, -id1 - 1, id2, -1, add_e, 2);
            }
            break;
        case 4:
            d = orc_d_pp(v1, v3);
            if(d < dHat)
            {
                FLAG01();
                emit(s, -id1 - 1, id3, -1, add_e, 2);
            }
            break;
        case 5:
            d = orc_d_pe(v1, v2, v3);
            if(d < dHat)
            {
                FLAG01();
                emit(s, -id1 - 1, id2, id3, add_e, 3);
            }
            break;
        case 6:
            d = orc_d_pe(v2, v0, v1);
            if(d < dHat)
            {
                FLAG23();
                emit(s, -id2 - 1, id0, id1, add_e, 3);
            }
            break;
        case 7:
            d = orc_d_pe(v3, v0, v1);
            if(d < dHat)
            {
                FLAG23();
                emit(s, -id3 - 1, id0, id1, add_e, 3);
            }
            break;
        case 8:
            d = orc_d_ee(v0, v1, v2, v3);
            if(d < dHat)
                emit(s, id0, id1, id2, id3, 4);
            break;
        default: break;
    }
#undef FLAG01
#undef FLAG23
}

static int body(const sgipc_scene_desc* s, int v) { return s->body_id ? s->body_id[v] : -1; }
static int bt(const sgipc_scene_desc* s, int v) { return s->btype ? s->btype[v] : 0; }

/* exclusion rules of _selfQuery_vf (mlbvh.cu:1343-1366) */
static int vf_ok(const sgipc_scene_desc* s, int v, const uint32_t* f)
{
    if(!((body(s, v) != body(s, (int)f[0])) || (body(s, v) == -1)))
        return 0;
    if(v == (int)f[0] || v == (int)f[1] || v == (int)f[2])
        return 0;
    if(bt(s, v) >= 2 && bt(s, f[0]) >= 2 && bt(s, f[1]) >= 2 && bt(s, f[2]) >= 2)
        return 0;
    return 1;
}

/* exclusion rules of _selfQuery_ee (mlbvh.cu:1553-1580), self < obj */
static int ee_ok(const sgipc_scene_desc* s, int self, int obj)
{
    const uint32_t* a = s->edges + 2 * self;
    const uint32_t* b = s->edges + 2 * obj;
    if(self == obj)
        return 0;
    if(!((body(s, a[0]) != body(s, b[0])) || (body(s, a[0]) == -1)))
        return 0;
    if(a[0] == b[0] || a[0] == b[1] || a[1] == b[0] || a[1] == b[1] || obj < self)
        return 0;
    if(bt(s, a[0]) >= 2 && bt(s, a[1]) >= 2 && bt(s, b[0]) >= 2 && bt(s, b[1]) >= 2)
        return 0;
    return 1;
}

int orc_contact_pairs_bruteforce(const sgipc_scene_desc* s, const double* x, double dhat2,
                                 int32_t* pairs, int cap, int32_t counts[5])
{
    sink_t sk = {pairs, cap, 0, counts};
    memset(counts, 0, sizeof(int32_t) * 5);
    const double* xr = s->x_rest ? s->x_rest : s->x;
    for(int i = 0; i < s->n_surf_verts; i++)
    {
        int v = (int)s->surf_verts[i];
        for(int f = 0; f < s->n_faces; f++)
        {
            const uint32_t* F = s->faces + 3 * f;
            if(vf_ok(s, v, F))
                check_pt(x, v, F[0], F[1], F[2], dhat2, &sk);
        }
    }
    for(int a = 0; a < s->n_edges; a++)
        for(int b = a + 1; b < s->n_edges; b++)
            if(ee_ok(s, a, b))
                check_ee(x, xr, s->edges[2 * a], s->edges[2 * a + 1], s->edges[2 * b],
                         s->edges[2 * b + 1], b, dhat2, &sk);
    return sk.n;
}

static int overlap(const double* l, const double* r, double gap)
{
    /* {upper xyz, lower xyz}; mlbvh.cu:31-40 */
    for(int c = 0; c < 3; c++)
    {
        if((r[3 + c] - l[c]) >= gap || (l[3 + c] - r[c]) >= gap)
            return 0;
    }
    return 1;
}

int orc_contact_pairs_bvh(const sgipc_scene_desc* s, const double* x, double dhat2,
                          int32_t* pairs, int cap, int32_t counts[5])
{
    sink_t sk = {pairs, cap, 0, counts};
    memset(counts, 0, sizeof(int32_t) * 5);
    const double* xr  = s->x_rest ? s->x_rest : s->x;
    double        gap = sqrt(dhat2);
    uint32_t      stack[64];
    if(s->n_faces > 0)
    {
        int       n     = s->n_faces;
        uint64_t* keys  = (uint64_t*)malloc(sizeof(uint64_t) * n);
        uint32_t* nodes = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (2 * n - 1));
        double*   bvs   = (double*)malloc(sizeof(double) * 6 * (2 * n - 1));
        orc_lbvh_build(x, s->faces, 3, n, keys, nodes, bvs);
        for(int i = 0; i < s->n_surf_verts; i++)
        {
            int    v     = (int)s->surf_verts[i];
            double bv[6] = {x[3 * v], x[3 * v + 1], x[3 * v + 2], x[3 * v], x[3 * v + 1],
                            x[3 * v + 2]};
            if(n == 1)
            {
                if(overlap(bv, bvs, gap) && vf_ok(s, v, s->faces))
                    check_pt(x, v, s->faces[0], s->faces[1], s->faces[2], dhat2, &sk);
                continue;
            }
            int sp      = 0;
            stack[sp++] = 0;
            do
            {
                uint32_t node = stack[--sp];
                uint32_t ch[2] = {nodes[4 * node + 1], nodes[4 * node + 2]};
                for(int c = 0; c < 2; c++)
                {
                    uint32_t C = ch[c];
                    if(!overlap(bv, bvs + 6 * C, gap))
                        continue;
                    uint32_t obj = nodes[4 * C + 3];
                    if(obj != 0xFFFFFFFFu)
                    {
                        const uint32_t* F = s->faces + 3 * obj;
                        if(vf_ok(s, v, F))
                            check_pt(x, v, F[0], F[1], F[2], dhat2, &sk);
                    }
                    else if(sp < 64)
                        stack[sp++] = C;
                }
            } while(sp > 0);
        }
        free(keys);
        free(nodes);
        free(bvs);
    }
    if(s->n_edges > 1)
    {
        int       n     = s->n_edges;
        uint64_t* keys  = (uint64_t*)malloc(sizeof(uint64_t) * n);
        uint32_t* nodes = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (2 * n - 1));
        double*   bvs   = (double*)malloc(sizeof(double) * 6 * (2 * n - 1));
        orc_lbvh_build(x, s->edges, 2, n, keys, nodes, bvs);
        for(int i = 0; i < n; i++)
        {
            int           leaf = i + n - 1;
            const double* bv   = bvs + 6 * leaf;
            int           self = (int)nodes[4 * leaf + 3];
            int           sp   = 0;
            stack[sp++]        = 0;
            do
            {
                uint32_t node  = stack[--sp];
                uint32_t ch[2] = {nodes[4 * node + 1], nodes[4 * node + 2]};
                for(int c = 0; c < 2; c++)
                {
                    uint32_t C = ch[c];
                    if(!overlap(bv, bvs + 6 * C, gap))
                        continue;
                    uint32_t obj = nodes[4 * C + 3];
                    if(obj != 0xFFFFFFFFu)
                    {
                        if(ee_ok(s, self, (int)obj))
                            check_ee(x, xr, s->edges[2 * self], s->edges[2 * self + 1],
                                     s->edges[2 * obj], s->edges[2 * obj + 1], (int)obj,
                                     dhat2, &sk);
                    }
                    else if(sp < 64)
                        stack[sp++] = C;
                }
            } while(sp > 0);
        }
        free(keys);
        free(nodes);
        free(bvs);
    }
    return sk.n;
}

/* ---------------------------------------------------------------- barrier */

/* decode a code into (type, ids): type 0 PP, 1 PE, 2 PT, 3 EE;
 * ids in the order the barrier uses them */
static int decode(const int32_t c[4], int32_t ids[4])
{
    if(c[0] >= 0)
    {
        ids[0] = c[0];
        ids[1] = c[1];
        ids[2] = c[2];
        ids[3] = c[3];
        return 3; /* EE (w >= 0; mollified EE never emitted) */
    }
    ids[0] = -c[0] - 1;
    if(c[2] < 0)
    {
        ids[1] = c[1];
        return 0; /* PP */
    }
    if(c[3] < 0)
    {
        ids[1] = c[1];
        ids[2] = c[2];
        return 1; /* PE */
    }
    ids[1] = c[1];
    ids[2] = c[2];
    ids[3] = c[3];
    return 2; /* PT */
}

static double type_d2(int type, const double* x, const int32_t* ids)
{
    switch(type)
    {
        case 0: return orc_d_pp(x + 3 * ids[0], x + 3 * ids[1]);
        case 1: return orc_d_pe(x + 3 * ids[0], x + 3 * ids[1], x + 3 * ids[2]);
        case 2: return orc_d_pt(x + 3 * ids[0], x + 3 * ids[1], x + 3 * ids[2], x + 3 * ids[3]);
        default: return orc_d_ee(x + 3 * ids[0], x + 3 * ids[1], x + 3 * ids[2], x + 3 * ids[3]);
    }
}

double orc_barrier_energy(const double* x, const double* xr, const int32_t code[4], double dhat2,
                          double kappa)
{
    int32_t ids[4];
    int     type = decode(code, ids);
    (void)xr;
    double dis  = type_d2(type, x, ids);
    double I5   = dis / dhat2;
    double lenE = dis - dhat2;
    return kappa * lenE * lenE * log(I5) * log(I5);
}

double orc_barrier_lambda0(double I5, double dhat2, double kappa)
{
    double L = log(I5);
    return -(4 * kappa * dhat2 * dhat2
             * (4 * I5 + L - 3 * I5 * I5 * L * L + 6 * I5 * L - 2 * I5 * I5 + I5 * L * L
                - 7 * I5 * I5 * L - 2))
           / I5;
}

/* gradient of the unsigned distance d w.r.t. the primitive's vertices */
static void dist_grad(int type, const double* x, const int32_t* ids, double* g)
{
    d3 x0 = ld(x + 3 * ids[0]), x1 = ld(x + 3 * ids[1]);
    if(type == 0)
    {
        d3     u = sub(x0, x1);
        double l = sqrt(sqn(u));
        g[0] = u.x / l, g[1] = u.y / l, g[2] = u.z / l;
        g[3] = -g[0], g[4] = -g[1], g[5] = -g[2];
        return;
    }
    d3 x2 = ld(x + 3 * ids[2]);
    if(type == 1)
    {
        d3     e = sub(x2, x1);
        double t = dot(sub(x0, x1), e) / sqn(e);
        d3     p = {x1.x + t * e.x, x1.y + t * e.y, x1.z + t * e.z};
        d3     u = sub(x0, p);
        double l = sqrt(sqn(u));
        u.x /= l, u.y /= l, u.z /= l;
        double w[3] = {1.0, -(1.0 - t), -t};
        for(int k = 0; k < 3; k++)
        {
            g[3 * k + 0] = w[k] * u.x;
            g[3 * k + 1] = w[k] * u.y;
            g[3 * k + 2] = w[k] * u.z;
        }
        return;
    }
    d3 x3 = ld(x + 3 * ids[3]);
    if(type == 2)
    {
        d3     e1 = sub(x2, x1), e2 = sub(x3, x1), q = sub(x0, x1);
        d3     nv = cross(e1, e2);
        double h  = dot(q, nv) / sqn(nv);
        d3     pq = {q.x - h * nv.x, q.y - h * nv.y, q.z - h * nv.z}; /* p - x1 */
        double a = sqn(e1), b = dot(e1, e2), c = sqn(e2);
        double r1 = dot(pq, e1), r2 = dot(pq, e2), D = a * c - b * b;
        double be = (c * r1 - b * r2) / D, ga = (a * r2 - b * r1) / D;
        d3     u  = {h * nv.x, h * nv.y, h * nv.z};
        double l  = sqrt(sqn(u));
        u.x /= l, u.y /= l, u.z /= l;
        double w[4] = {1.0, -(1.0 - be - ga), -be, -ga};
        for(int k = 0; k < 4; k++)
        {
            g[3 * k + 0] = w[k] * u.x;
            g[3 * k + 1] = w[k] * u.y;
            g[3 * k + 2] = w[k] * u.z;
        }
        return;
    }
    {
        d3     u = sub(x1, x0), v = sub(x3, x2), w = sub(x0, x2);
        double a = sqn(u), b = dot(u, v), c = sqn(v), d = dot(u, w), e = dot(v, w);
        double D = a * c - b * b;
        double s = (b * e - c * d) / D, t = (a * e - b * d) / D;
        d3     pq = {w.x + s * u.x - t * v.x, w.y + s * u.y - t * v.y, w.z + s * u.z - t * v.z};
        double l  = sqrt(sqn(pq));
        pq.x /= l, pq.y /= l, pq.z /= l;
        double wt[4] = {1.0 - s, s, -(1.0 - t), -t};
        for(int k = 0; k < 4; k++)
        {
            g[3 * k + 0] = wt[k] * pq.x;
            g[3 * k + 1] = wt[k] * pq.y;
            g[3 * k + 2] = wt[k] * pq.z;
        }
    }
}

int orc_barrier_grad_hess(const double* x, const double* xr, const int32_t code[4], double dhat2,
                          double kappa, int32_t ids[4], double* g, double* H)
{
    (void)xr;
    int    type = decode(code, ids);
    int    nv   = (type == 0) ? 2 : (type == 1) ? 3 : 4;
    double dis  = type_d2(type, x, ids);
    double d    = sqrt(dis);
    double dhs  = sqrt(dhat2);
    double I5   = pow(d / dhs, 2);
    double s    = d / dhs;
    double L    = log(I5);
    double pk1  = s * 2 * (2 * kappa * dhat2 * dhat2 * L * (I5 - 1) * (I5 + I5 * L - 1)) / I5;
    double lam0 = orc_barrier_lambda0(I5, dhat2, kappa);
    if(d * d < 1e-6 * dhat2)
        lam0 = orc_barrier_lambda0(1e-6, dhat2, kappa);
    double gd[12];
    dist_grad(type, x, ids, gd);
    int n = 3 * nv;
    for(int i = 0; i < n; i++)
        gd[i] /= dhs; /* ds/dx */
    for(int i = 0; i < n; i++)
        g[i] = pk1 * gd[i];
    for(int i = 0; i < n; i++)
        for(int j = 0; j < n; j++)
            H[i * n + j] = lam0 * gd[i] * gd[j];
    return nv;
}
