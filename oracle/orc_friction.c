/*
 * orc_friction.c -- lagged IPC friction (self contacts and ground).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates StiffGIPC (USE_FRICTION, RANK 2, SFCLAMPING_ORDER 1):
 *   lagged normal force, closest-point coordinates and tangent bases at the
 *   start of a time step:  _calFrictionLastH_DistAndTan GIPC.cu:8330-8441,
 *                          _calFrictionLastH_gd :8304-8328,
 *                          buildFrictionSets :8601-8637
 *   tangent bases / closest points / relative displacements / T matrices:
 *                          FrictionUtils.cuh:52-420
 *   C1 static-friction mollifier f0, f1/x, f2: FrictionUtils.cuh:419-499
 *   gradient  _calFrictionGradient :5089-5273, _calFrictionGradient_gd :5049-5088
 *   Hessian   _calFrictionHessian :1441-1804 (2x2 PSD projection,
 *             makePDGeneral :33-59), _calFrictionHessian_gd :1359-1440
 *   energy    __cal_Friction_energy :1279-1352, __cal_Friction_gd_energy :1254-1276
 * eps^2 = fDhat dt^2 with fDhat = 1e-4 bboxDiag^2 (GIPC.cu:8544, :9128).
 *
 * Pairs use the signed int4 contact code; only the non-mollified kinds carry
 * friction (EE with w >= 0, PP/PE with y >= 0, PT).  The ground Hessian keeps
 * the reference's x/z tangent plane (it assumes the normal (0, 1, 0)).
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

typedef struct
{
    double x, y, z;
} f3;

static f3 L3(const double* p)
{
    f3 r = {p[0], p[1], p[2]};
    return r;
}
static f3 sub(f3 a, f3 b)
{
    f3 r = {a.x - b.x, a.y - b.y, a.z - b.z};
    return r;
}
static f3 add(f3 a, f3 b)
{
    f3 r = {a.x + b.x, a.y + b.y, a.z + b.z};
    return r;
}
static f3 mul(f3 a, double s)
{
    f3 r = {a.x * s, a.y * s, a.z * s};
    return r;
}
static double dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static double sqn(f3 a) { return dot(a, a); }
static f3     cross(f3 a, f3 b)
{
    f3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
static f3 nrm(f3 a) { return mul(a, 1.0 / sqrt(sqn(a))); }

/* basis: 3x2 row-major, columns b0 (basis[0], [2], [4]) and b1 ([1], [3], [5]) */
static void set_basis(double* B, f3 c0, f3 c1)
{
    B[0] = c0.x, B[2] = c0.y, B[4] = c0.z;
    B[1] = c1.x, B[3] = c1.y, B[5] = c1.z;
}

/* 2x2 inverse through the adjugate */
static void inv2(const double m[4], double o[4])
{
    double det = m[0] * m[3] - m[1] * m[2];
    o[0]       = m[3] / det;
    o[1]       = -m[1] / det;
    o[2]       = -m[2] / det;
    o[3]       = m[0] / det;
}

/* FrictionUtils.cuh:52-61 */
static void basis_pt(f3 v0, f3 v1, f3 v2, f3 v3, double* B)
{
    f3 v12 = sub(v2, v1);
    f3 c   = nrm(cross(cross(v12, sub(v3, v1)), v12));
    (void)v0;
    set_basis(B, nrm(v12), c);
}
/* :62-82 */
static void coord_pt(f3 v0, f3 v1, f3 v2, f3 v3, double beta[2])
{
    f3     v12 = sub(v2, v1), v13 = sub(v3, v1), v10 = sub(v0, v1);
    double m[4] = {sqn(v12), dot(v12, v13), dot(v13, v12), sqn(v13)}, mi[4];
    inv2(m, mi);
    double b0 = dot(v12, v10), b1 = dot(v13, v10);
    beta[0]   = mi[0] * b0 + mi[1] * b1;
    beta[1]   = mi[2] * b0 + mi[3] * b1;
}
/* :161-173 */
static void basis_ee(f3 v0, f3 v1, f3 v2, f3 v3, double* B)
{
    f3 v01 = sub(v1, v0);
    f3 c   = nrm(cross(cross(v01, sub(v3, v2)), v01));
    set_basis(B, nrm(v01), c);
}
/* :174-194 */
static void coord_ee(f3 v0, f3 v1, f3 v2, f3 v3, double g[2])
{
    f3     e20 = sub(v0, v2), e01 = sub(v1, v0), e23 = sub(v3, v2);
    double m[4] = {sqn(e01), -dot(e23, e01), -dot(e23, e01), sqn(e23)}, mi[4];
    inv2(m, mi);
    double r0 = -dot(e20, e01), r1 = dot(e20, e23);
    g[0]      = mi[0] * r0 + mi[1] * r1;
    g[1]      = mi[2] * r0 + mi[3] * r1;
}
/* :266-278 */
static void basis_pe(f3 v0, f3 v1, f3 v2, double* B)
{
    f3 v12 = sub(v2, v1);
    f3 c   = nrm(cross(v12, sub(v0, v1)));
    set_basis(B, nrm(v12), c);
}
/* :279-285 */
static double coord_pe(f3 v0, f3 v1, f3 v2)
{
    f3 e12 = sub(v2, v1);
    return dot(sub(v0, v1), e12) / sqn(e12);
}
/* :342-372 */
static void basis_pp(f3 v0, f3 v1, double* B)
{
    f3 v01 = sub(v1, v0);
    f3 xc  = {0.0, -v01.z, v01.y};
    f3 yc  = {v01.z, 0.0, -v01.x};
    if(sqn(xc) > sqn(yc))
        set_basis(B, nrm(xc), nrm(cross(v01, xc)));
    else
        set_basis(B, nrm(yc), nrm(cross(v01, yc)));
}

/* per-kind vertex weights w_k such that the relative tangential displacement
 * is sum_k w_k dx_k and T = [w_k basis] (computeRelDX_* / computeT_*) */
static int weights(int kind, const double c[2], double w[4])
{
    switch(kind)
    {
        case 0: /* PP */
            w[0] = 1.0, w[1] = -1.0;
            return 2;
        case 1: /* PE */
            w[0] = 1.0, w[1] = c[0] - 1.0, w[2] = -c[0];
            return 3;
        case 2: /* PT */
            w[0] = 1.0, w[1] = -1.0 + c[0] + c[1], w[2] = -c[0], w[3] = -c[1];
            return 4;
        default: /* EE */
            w[0] = 1.0 - c[0], w[1] = c[0], w[2] = c[1] - 1.0, w[3] = -c[1];
            return 4;
    }
}

/* decode a contact code into kind (0 PP, 1 PE, 2 PT, 3 EE, -1 no friction)
 * and vertex ids (_calFrictionLastH_DistAndTan's filter) */
static int friction_kind(const int32_t* c, int ids[4])
{
    if(c[0] >= 0)
    {
        if(c[3] < 0)
            return -1;
        ids[0] = c[0], ids[1] = c[1], ids[2] = c[2], ids[3] = c[3];
        return 3;
    }
    ids[0] = -c[0] - 1;
    ids[1] = c[1];
    if(c[2] < 0)
        return c[1] >= 0 ? 0 : -1;
    ids[2] = c[2];
    if(c[3] < 0)
        return c[1] >= 0 ? 1 : -1;
    ids[3] = c[3];
    return 2;
}

int orc_friction_lag(const double* x, const int32_t* code, double dhat2, double kappa,
                     double* lam, double coord[2], double basis[6])
{
    int ids[4];
    int kind = friction_kind(code, ids);
    if(kind < 0)
        return -1;
    f3     v0 = L3(x + 3 * ids[0]), v1 = L3(x + 3 * ids[1]);
    double dis;
    coord[0] = coord[1] = 0.0;
    if(kind == 0)
    {
        dis = orc_d_pp(x + 3 * ids[0], x + 3 * ids[1]);
        basis_pp(v0, v1, basis);
    }
    else if(kind == 1)
    {
        f3 v2 = L3(x + 3 * ids[2]);
        dis   = orc_d_pe(x + 3 * ids[0], x + 3 * ids[1], x + 3 * ids[2]);
        coord[0] = coord_pe(v0, v1, v2);
        basis_pe(v0, v1, v2, basis);
    }
    else
    {
        f3 v2 = L3(x + 3 * ids[2]), v3 = L3(x + 3 * ids[3]);
        if(kind == 2)
        {
            dis = orc_d_pt(x + 3 * ids[0], x + 3 * ids[1], x + 3 * ids[2], x + 3 * ids[3]);
            coord_pt(v0, v1, v2, v3, coord);
            basis_pt(v0, v1, v2, v3, basis);
        }
        else
        {
            dis = orc_d_ee(x + 3 * ids[0], x + 3 * ids[1], x + 3 * ids[2], x + 3 * ids[3]);
            coord_ee(v0, v1, v2, v3, coord);
            basis_ee(v0, v1, v2, v3, basis);
        }
    }
    double L = log(dis / dhat2);
    *lam     = -kappa * 2.0 * sqrt(dis)
           * (L * L * (2 * dis - 2 * dhat2) + (2 * L * (dis - dhat2) * (dis - dhat2)) / dis);
    return kind;
}

double orc_friction_lag_ground(const double* x, const double n[3], double off, double dhat2,
                               double kappa)
{
    double dist  = n[0] * x[0] + n[1] * x[1] + n[2] * x[2] - off;
    double dist2 = dist * dist;
    double t     = dist2 - dhat2;
    double g_b   = t * log(dist2 / dhat2) * -2.0 - (t * t) / dist2;
    return -kappa * 2.0 * sqrt(dist2) * g_b;
}

/* tangential relative displacement u = basis^T sum_k w_k (x_k - xprev_k) */
static int rel_u(const double* x, const double* xp, const int32_t* code, const double c[2],
                 const double B[6], int ids[4], double w[4], double u[2])
{
    int kind = friction_kind(code, ids);
    int m    = weights(kind, c, w);
    f3  r    = {0, 0, 0};
    for(int k = 0; k < m; k++)
        r = add(r, mul(sub(L3(x + 3 * ids[k]), L3(xp + 3 * ids[k])), w[k]));
    u[0] = B[0] * r.x + B[2] * r.y + B[4] * r.z;
    u[1] = B[1] * r.x + B[3] * r.y + B[5] * r.z;
    return m;
}

/* C1 mollifier, x2 = |u|^2 (FrictionUtils.cuh:450-465) */
static double f0_c1(double x2, double eps) { return x2 * (-sqrt(x2) / 3.0 + eps) / (eps * eps) + eps / 3.0; }
static double f1_div_c1(double x2, double eps) { return (-sqrt(x2) + 2.0 * eps) / (eps * eps); }
static double f2_c1(double x2, double eps) { return 2.0 * (eps - sqrt(x2)) / (eps * eps); }

double orc_friction_energy(const double* x, const double* xp, const int32_t* code,
                           const double coord[2], const double basis[6], double lam, double eps2)
{
    int    ids[4];
    double w[4], u[2];
    rel_u(x, xp, code, coord, basis, ids, w, u);
    double s2 = u[0] * u[0] + u[1] * u[1];
    if(s2 > eps2)
        return lam * sqrt(s2);
    return lam * f0_c1(s2, sqrt(eps2));
}

/* symmetric 2x2 PSD projection by eigenvalue clamping, unchanged when both
 * eigenvalues are non-negative (makePDGeneral) */
static void psd2(double M[4])
{
    double a = M[0], b = 0.5 * (M[1] + M[2]), d = M[3];
    double tr = 0.5 * (a + d), df = 0.5 * (a - d);
    double r  = sqrt(df * df + b * b);
    double l0 = tr - r, l1 = tr + r;
    if(l0 >= 0.0)
        return;
    /* eigenvector of l1 */
    double vx, vy;
    if(r == 0.0)
        vx = 1.0, vy = 0.0;
    else if(df >= 0.0)
    {
        vx = df + r, vy = b;
    }
    else
    {
        vx = b, vy = r - df;
    }
    double nv = sqrt(vx * vx + vy * vy);
    vx /= nv, vy /= nv;
    double k = l1 > 0.0 ? l1 : 0.0;
    M[0]     = k * vx * vx;
    M[1] = M[2] = k * vx * vy;
    M[3]        = k * vy * vy;
}

int orc_friction_grad_hess(const double* x, const double* xp, const int32_t* code,
                           const double coord[2], const double basis[6], double lam, double eps2,
                           double coef, int32_t ids_out[4], double g[12], double H[144])
{
    int    ids[4];
    double w[4], u[2];
    int    m   = rel_u(x, xp, code, coord, basis, ids, w, u);
    double eps = sqrt(eps2);
    double s2  = u[0] * u[0] + u[1] * u[1], s = sqrt(s2);
    double f1  = (s2 > eps2) ? 1.0 / s : f1_div_c1(s2, eps);
    double q[2] = {u[0] * f1, u[1] * f1};
    /* gradient: coef lam T q with T rows w_k basis */
    for(int k = 0; k < m; k++)
        for(int c = 0; c < 3; c++)
            g[3 * k + c] = coef * lam * w[k] * (basis[2 * c] * q[0] + basis[2 * c + 1] * q[1]);
    /* Hessian: coef lam T PSD(M2) T^T */
    double M2[4];
    if(s2 > eps2)
    {
        double c3 = 1.0 / (s2 * s);
        M2[0]     = 1.0 / s - u[0] * u[0] * c3;
        M2[1]     = -u[0] * u[1] * c3;
        M2[2]     = -u[1] * u[0] * c3;
        M2[3]     = 1.0 / s - u[1] * u[1] * c3;
    }
    else
    {
        double f2 = f2_c1(s2, eps);
        if(f2 != f1 && s2 != 0.0)
        {
            double c2 = (f1 - f2) / s2;
            M2[0]     = f1 - u[0] * u[0] * c2;
            M2[1]     = -u[0] * u[1] * c2;
            M2[2]     = -u[1] * u[0] * c2;
            M2[3]     = f1 - u[1] * u[1] * c2;
        }
        else
        {
            M2[0] = f1, M2[1] = 0.0, M2[2] = 0.0, M2[3] = f1;
        }
    }
    psd2(M2);
    int n = 3 * m;
    for(int p = 0; p < n; p++)
    {
        int    kp = p / 3, cp = p % 3;
        double tp0 = w[kp] * basis[2 * cp], tp1 = w[kp] * basis[2 * cp + 1];
        double a0 = tp0 * M2[0] + tp1 * M2[2], a1 = tp0 * M2[1] + tp1 * M2[3];
        for(int q2 = 0; q2 < n; q2++)
        {
            int    kq = q2 / 3, cq = q2 % 3;
            double tq0 = w[kq] * basis[2 * cq], tq1 = w[kq] * basis[2 * cq + 1];
            H[p * n + q2] = coef * lam * (a0 * tq0 + a1 * tq1);
        }
    }
    for(int k = 0; k < m; k++)
        ids_out[k] = ids[k];
    return m;
}

/* ---------------------------------------------------------------- ground */
static void vproj(const double* x, const double* xp, const double n[3], double v[3])
{
    double d[3] = {x[0] - xp[0], x[1] - xp[1], x[2] - xp[2]};
    double dn   = d[0] * n[0] + d[1] * n[1] + d[2] * n[2];
    for(int c = 0; c < 3; c++)
        v[c] = d[c] - n[c] * dn;
}

double orc_friction_ground_energy(const double* x, const double* xp, const double n[3],
                                  double lam, double eps2)
{
    double v[3];
    vproj(x, xp, n, v);
    double m2  = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    double eps = sqrt(eps2);
    if(m2 > eps * eps)
        return lam * (sqrt(m2) - eps * 0.5);
    return lam * m2 / eps * 0.5;
}

void orc_friction_ground_grad_hess(const double* x, const double* xp, const double n[3],
                                   double lam, double eps2, double coef, double g[3], double H[9])
{
    double v[3];
    vproj(x, xp, n, v);
    double m2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    double eps = sqrt(eps2), mult = coef * lam;
    double s   = (m2 > eps2) ? mult / sqrt(m2) : mult / eps;
    for(int c = 0; c < 3; c++)
        g[c] = v[c] * s;
    memset(H, 0, sizeof(double) * 9);
    if(m2 > eps2)
    {
        double mag = sqrt(m2);
        double M[4] = {v[0] * v[0] * -mult / m2 / mag + mult / mag, v[0] * v[2] * -mult / m2 / mag,
                       v[0] * v[2] * -mult / m2 / mag, v[2] * v[2] * -mult / m2 / mag + mult / mag};
        /* reconstruct from the positive eigenpairs (__makePD2x2 + sum) */
        double a = M[0], b = M[1], d = M[3];
        double tr = 0.5 * (a + d), df = 0.5 * (a - d), r = sqrt(df * df + b * b);
        double P[4] = {0, 0, 0, 0};
        for(int e = 0; e < 2; e++)
        {
            double l = e ? tr + r : tr - r;
            if(!(l > 0.0))
                continue;
            double vx, vy;
            if(r == 0.0)
                vx = e ? 1.0 : 0.0, vy = e ? 0.0 : 1.0;
            else if((df >= 0.0) == (e == 1))
                vx = df + (e ? r : -r), vy = b;
            else
                vx = b, vy = (e ? r : -r) - df;
            double nv = sqrt(vx * vx + vy * vy);
            if(nv == 0.0)
                continue;
            vx /= nv, vy /= nv;
            P[0] += l * vx * vx;
            P[1] += l * vx * vy;
            P[2] += l * vx * vy;
            P[3] += l * vy * vy;
        }
        H[0] = P[0], H[2] = P[1], H[6] = P[2], H[8] = P[3];
    }
    else
    {
        H[0] = mult / eps;
        H[8] = mult / eps;
    }
}
