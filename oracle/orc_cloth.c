/*
 * orc_cloth.c -- cloth: Baraff-Witkin membrane and discrete-shell bending.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates StiffGIPC/femEnergy.cu:
 *   __calculateDm2D_double            :262-309 (rest frame by the minimal
 *                                      rotation taking the normal to +z)
 *   __computePEPF_BaraffWitkinStretch_double :360-412
 *   __project_BaraffWitkinStretch_H   :565-663
 *   __project_BaraffWitkinShear_H     :666-740
 *   __cal_BaraffWitkinStretch_energy  :1003-1041
 *   _calculate_triangle_fem_gradient_hessian :2215-2300 (F = Ds Dm^-1,
 *                                      g = area dt^2 dF/dx^T vec P)
 *   _calculate_bending_gradient_hessian :2431-2566 (H = 2((t - t0) d2t + dt dt^T)
 *                                      made PSD, scaled by k |e0| dt^2)
 *   __cal_bending_energy              :2590-2630 (k |e0| (t - t0)^2)
 *   LibShell::edgeTheta / angle       :102-260 (signed dihedral angle)
 * and StiffGIPC/main.cu:303-307 (stretch = E/(2(1+nu)),
 * bend = E_b h^3 / (24 (1 - nu^2)), shear = 0.03 stretch strainRate),
 * StiffGIPC/load_mesh.cpp:320-379 (bending hinges: interior edges (a < b) in
 * increasing order, opposite vertices in triangle order).
 *
 * The membrane Hessian is formed here as the exact Hessian of each term
 * (stretch along u, stretch along v, shear) in vec(F) space and projected by
 * the generic Jacobi eigen-solver; the reference's closed-form projections
 * are the exact PSD projections of the same terms, so the two agree to
 * rounding without sharing an eigenbasis.
 * The dihedral angle is theta = atan2(-V |e1|, (e1.e2)(e1.e3) - |e1|^2 (e2.e3))
 * with e_k = q_k - q_0 and V = e1.(e2 x e3) -- the same angle as LibShell's
 * 2 atan2((n0 x n1).a^, n0.n1 + |n0||n1|) with n0 = (q0-q2)x(q1-q2),
 * n1 = (q1-q3)x(q0-q3), a = q1 - q0 -- differentiated through the chain rule
 * of atan2(u, w).  The 12x12 bending Hessian is projected by the generic
 * eigen-solver (makePDSNK<double, 12>, femEnergy.cu:64-84).
 *
 * Cloth set-up that the reference leaves out of initFEM (main.cu:351-352):
 * area = rest triangle area, lumped mass area * thickness * density / 3 per
 * vertex.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

void orc_cloth_params(double E, double nu, double bend_E, double thickness, double strain_rate,
                      double* stretch, double* shear, double* bend)
{
    *stretch = E / (2.0 * (1.0 + nu));
    *bend    = bend_E * thickness * thickness * thickness / (24.0 * (1.0 - nu * nu));
    *shear   = 0.03 * (*stretch) * strain_rate;
}

void orc_tri_precompute(const double* xr, const uint32_t t[3], double dminv[4], double* area)
{
    const double* a = xr + 3 * t[0];
    const double* b = xr + 3 * t[1];
    const double* c = xr + 3 * t[2];
    double        u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double        v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double        n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                          u[0] * v[1] - u[1] * v[0]};
    double        nn   = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    *area              = 0.5 * nn;
    for(int k = 0; k < 3; k++)
        n[k] /= nn;
    /* rotation taking n to +z: R = I + [r]x + [r]x^2 / (1 + cos), r = n x z */
    double r[3] = {n[1], -n[0], 0.0};
    double cs   = n[2];
    double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if(cs + 1.0 == 0.0)
    {
        R[0] = -1.0;
        R[4] = -1.0;
    }
    else
    {
        double K[9] = {0, -r[2], r[1], r[2], 0, -r[0], -r[1], r[0], 0};
        double K2[9];
        for(int i = 0; i < 3; i++)
            for(int j = 0; j < 3; j++)
                K2[i * 3 + j] = K[i * 3 + 0] * K[0 * 3 + j] + K[i * 3 + 1] * K[1 * 3 + j]
                                + K[i * 3 + 2] * K[2 * 3 + j];
        for(int k = 0; k < 9; k++)
            R[k] = R[k] + (K[k] + K2[k] * (1.0 / (1.0 + cs)));
    }
    double p[3][2];
    const double* P[3] = {a, b, c};
    for(int k = 0; k < 3; k++)
        for(int i = 0; i < 2; i++)
            p[k][i] = R[i * 3 + 0] * P[k][0] + R[i * 3 + 1] * P[k][1] + R[i * 3 + 2] * P[k][2];
    /* Dm = [p1 - p0, p2 - p0] (columns), row-major 2x2 */
    double D[4] = {p[1][0] - p[0][0], p[2][0] - p[0][0], p[1][1] - p[0][1], p[2][1] - p[0][1]};
    double det  = D[0] * D[3] - D[1] * D[2];
    dminv[0]    = D[3] / det;
    dminv[1]    = -D[1] / det;
    dminv[2]    = -D[2] / det;
    dminv[3]    = D[0] / det;
}

/* stretch term along one material axis: f(s) = (s-1)^2 + [s>1] r (s-1)^3 */
static void stretch_f(double s, double r, double* f, double* f1, double* f2)
{
    double d = s - 1.0;
    int    c = (s * s > 1.0);
    if(f)
        *f = d * d + (c ? r * d * d * d : 0.0);
    if(f1)
        *f1 = 2.0 * d + (c ? 3.0 * r * d * d : 0.0);
    if(f2)
        *f2 = 2.0 + (c ? 6.0 * r * d : 0.0);
}

/* F is 3x2 row-major: F[2 i + c] */
double orc_bw_psi(const double F[6], double stretch, double shear, double r)
{
    double su = sqrt(F[0] * F[0] + F[2] * F[2] + F[4] * F[4]);
    double sv = sqrt(F[1] * F[1] + F[3] * F[3] + F[5] * F[5]);
    double I6 = F[0] * F[1] + F[2] * F[3] + F[4] * F[5];
    double fu, fv;
    stretch_f(su, r, &fu, NULL, NULL);
    stretch_f(sv, r, &fv, NULL, NULL);
    return stretch * (fu + fv) + shear * I6 * I6;
}

void orc_bw_pk1(const double F[6], double stretch, double shear, double r, double P[6])
{
    double su = sqrt(F[0] * F[0] + F[2] * F[2] + F[4] * F[4]);
    double sv = sqrt(F[1] * F[1] + F[3] * F[3] + F[5] * F[5]);
    double I6 = F[0] * F[1] + F[2] * F[3] + F[4] * F[5];
    double du, dv;
    stretch_f(su, r, NULL, &du, NULL);
    stretch_f(sv, r, NULL, &dv, NULL);
    for(int i = 0; i < 3; i++)
    {
        P[2 * i]     = stretch * du / su * F[2 * i] + 2.0 * shear * I6 * F[2 * i + 1];
        P[2 * i + 1] = stretch * dv / sv * F[2 * i + 1] + 2.0 * shear * I6 * F[2 * i];
    }
}

/* 6x6 Hessian in vec(F) = (F[:,0], F[:,1]) order (index i + 3 c) */
void orc_bw_hess(const double F[6], double stretch, double shear, double r, int project,
                 double H[36])
{
    memset(H, 0, sizeof(double) * 36);
    /* stretch: f'/s (I - n n^T) + f'' n n^T per axis */
    for(int c = 0; c < 2; c++)
    {
        double f[3] = {F[c], F[2 + c], F[4 + c]};
        double s    = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        double d1, d2, B[9];
        stretch_f(s, r, NULL, &d1, &d2);
        for(int i = 0; i < 3; i++)
            for(int j = 0; j < 3; j++)
            {
                double nn    = f[i] * f[j] / (s * s);
                B[i * 3 + j] = d1 / s * ((i == j) - nn) + d2 * nn;
            }
        if(project)
            orc_psd_project(3, B);
        for(int i = 0; i < 3; i++)
            for(int j = 0; j < 3; j++)
                H[(3 * c + i) * 6 + 3 * c + j] += stretch * B[i * 3 + j];
    }
    /* shear: psi = I6^2, d2 = 2 (g g^T + I6 S), g = (F[:,1], F[:,0]),
     * S swaps the two column blocks */
    double g[6] = {F[1], F[3], F[5], F[0], F[2], F[4]};
    double I6   = F[0] * F[1] + F[2] * F[3] + F[4] * F[5];
    double S[36];
    for(int a = 0; a < 6; a++)
        for(int b = 0; b < 6; b++)
            S[a * 6 + b] = 2.0 * (g[a] * g[b] + ((a % 3 == b % 3 && a / 3 != b / 3) ? I6 : 0.0));
    if(project)
        orc_psd_project(6, S);
    for(int k = 0; k < 36; k++)
        H[k] += shear * S[k];
}

static void tri_F(const double* x, const uint32_t t[3], const double dminv[4], double F[6])
{
    for(int i = 0; i < 3; i++)
    {
        double d1 = x[3 * t[1] + i] - x[3 * t[0] + i];
        double d2 = x[3 * t[2] + i] - x[3 * t[0] + i];
        F[2 * i]     = d1 * dminv[0] + d2 * dminv[2];
        F[2 * i + 1] = d1 * dminv[1] + d2 * dminv[3];
    }
}

double orc_tri_energy(const double* x, const uint32_t t[3], const double dminv[4], double area,
                      double stretch, double shear, double r)
{
    double F[6];
    tri_F(x, t, dminv, F);
    return area * orc_bw_psi(F, stretch, shear, r);
}

void orc_tri_grad_hess(const double* x, const uint32_t t[3], const double dminv[4], double area,
                       double stretch, double shear, double r, double dt, double g[9],
                       double H[81])
{
    double F[6], P[6], Hf[36];
    tri_F(x, t, dminv, F);
    orc_bw_pk1(F, stretch, shear, r, P);
    orc_bw_hess(F, stretch, shear, r, 1, Hf);
    /* dF[:,c]/dx_a = B[a][c] I */
    double B[3][2] = {{-(dminv[0] + dminv[2]), -(dminv[1] + dminv[3])},
                      {dminv[0], dminv[1]},
                      {dminv[2], dminv[3]}};
    double s       = area * dt * dt;
    for(int a = 0; a < 3; a++)
        for(int i = 0; i < 3; i++)
            g[3 * a + i] = s * (B[a][0] * P[2 * i] + B[a][1] * P[2 * i + 1]);
    for(int a = 0; a < 3; a++)
        for(int b = 0; b < 3; b++)
            for(int i = 0; i < 3; i++)
                for(int j = 0; j < 3; j++)
                {
                    double h = 0.0;
                    for(int c = 0; c < 2; c++)
                        for(int d = 0; d < 2; d++)
                            h += B[a][c] * B[b][d] * Hf[(3 * c + i) * 6 + 3 * d + j];
                    H[(3 * a + i) * 9 + 3 * b + j] = s * h;
                }
}

/* ------------------------------------------------------------- bending */
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void   cross3(const double* a, const double* b, double* c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
/* M = s [a]x added into the 3x3 block (bi, bj) of a 9x9 matrix */
static void add_skew(double* H9, int bi, int bj, const double* a, double s)
{
    double K[9] = {0, -a[2], a[1], a[2], 0, -a[0], -a[1], a[0], 0};
    for(int i = 0; i < 3; i++)
        for(int j = 0; j < 3; j++)
            H9[(3 * bi + i) * 9 + 3 * bj + j] += s * K[i * 3 + j];
}
static void add_eye(double* H9, int bi, int bj, double s)
{
    for(int i = 0; i < 3; i++)
        H9[(3 * bi + i) * 9 + 3 * bj + i] += s;
}

double orc_dihedral(const double* q0, const double* q1, const double* q2, const double* q3,
                    double g[12], double H[144])
{
    double e1[3], e2[3], e3[3];
    for(int i = 0; i < 3; i++)
    {
        e1[i] = q1[i] - q0[i];
        e2[i] = q2[i] - q0[i];
        e3[i] = q3[i] - q0[i];
    }
    double c23[3], c31[3], c12[3];
    cross3(e2, e3, c23);
    cross3(e3, e1, c31);
    cross3(e1, e2, c12);
    double V = dot3(e1, c23);
    double L = sqrt(dot3(e1, e1));
    double p = dot3(e1, e2), q = dot3(e1, e3), s = dot3(e1, e1), t = dot3(e2, e3);
    double u = -V * L, w = p * q - s * t;
    double theta = atan2(u, w);
    if(!g && !H)
        return theta;
    /* first derivatives over (e1, e2, e3) */
    double gV[9], gL[9] = {0}, gp[9] = {0}, gq[9] = {0}, gs[9] = {0}, gt[9] = {0};
    for(int i = 0; i < 3; i++)
    {
        gV[i]     = c23[i];
        gV[3 + i] = c31[i];
        gV[6 + i] = c12[i];
        gL[i]     = e1[i] / L;
        gp[i]     = e2[i];
        gp[3 + i] = e1[i];
        gq[i]     = e3[i];
        gq[6 + i] = e1[i];
        gs[i]     = 2.0 * e1[i];
        gt[3 + i] = e3[i];
        gt[6 + i] = e2[i];
    }
    double gu[9], gw[9], g9[9];
    double r2 = u * u + w * w;
    for(int k = 0; k < 9; k++)
    {
        gu[k] = -(L * gV[k] + V * gL[k]);
        gw[k] = q * gp[k] + p * gq[k] - t * gs[k] - s * gt[k];
        g9[k] = (w * gu[k] - u * gw[k]) / r2;
    }
    if(g)
        for(int i = 0; i < 3; i++)
        {
            g[i]     = -(g9[i] + g9[3 + i] + g9[6 + i]);
            g[3 + i] = g9[i];
            g[6 + i] = g9[3 + i];
            g[9 + i] = g9[6 + i];
        }
    if(!H)
        return theta;
    /* second derivatives */
    double HV[81] = {0}, HL[81] = {0}, Hw[81] = {0}, Hu[81], H9[81];
    add_skew(HV, 0, 1, e3, -1.0);
    add_skew(HV, 1, 0, e3, 1.0);
    add_skew(HV, 0, 2, e2, 1.0);
    add_skew(HV, 2, 0, e2, -1.0);
    add_skew(HV, 1, 2, e1, -1.0);
    add_skew(HV, 2, 1, e1, 1.0);
    for(int i = 0; i < 3; i++)
        for(int j = 0; j < 3; j++)
            HL[i * 9 + j] = ((i == j) - e1[i] * e1[j] / s) / L;
    add_eye(Hw, 0, 1, q);
    add_eye(Hw, 1, 0, q);
    add_eye(Hw, 0, 2, p);
    add_eye(Hw, 2, 0, p);
    add_eye(Hw, 0, 0, -2.0 * t);
    add_eye(Hw, 1, 2, -s);
    add_eye(Hw, 2, 1, -s);
    for(int a = 0; a < 9; a++)
        for(int b = 0; b < 9; b++)
        {
            Hw[a * 9 + b] += gp[a] * gq[b] + gq[a] * gp[b] - gs[a] * gt[b] - gt[a] * gs[b];
            Hu[a * 9 + b] = -(L * HV[a * 9 + b] + gV[a] * gL[b] + gL[a] * gV[b] + V * HL[a * 9 + b]);
        }
    double r4 = r2 * r2;
    for(int a = 0; a < 9; a++)
        for(int b = 0; b < 9; b++)
            H9[a * 9 + b] = (w * Hu[a * 9 + b] - u * Hw[a * 9 + b]) / r2
                            + (-2.0 * u * w * gu[a] * gu[b] + 2.0 * u * w * gw[a] * gw[b]
                               + (u * u - w * w) * (gu[a] * gw[b] + gw[a] * gu[b]))
                                  / r4;
    /* q0 = -(e1 + e2 + e3) direction: expand to the 4 vertices */
    for(int A = 0; A < 4; A++)
        for(int B = 0; B < 4; B++)
            for(int i = 0; i < 3; i++)
                for(int j = 0; j < 3; j++)
                {
                    double h = 0.0;
                    for(int a = 0; a < 3; a++)
                        for(int b = 0; b < 3; b++)
                        {
                            double ca = (A == 0) ? -1.0 : (A == a + 1 ? 1.0 : 0.0);
                            double cb = (B == 0) ? -1.0 : (B == b + 1 ? 1.0 : 0.0);
                            if(ca != 0.0 && cb != 0.0)
                                h += ca * cb * H9[(3 * a + i) * 9 + 3 * b + j];
                        }
                    H[(3 * A + i) * 12 + 3 * B + j] = h;
                }
    return theta;
}

double orc_bend_energy(const double* x, const uint32_t hinge[4], double theta0, double len0,
                       double bend)
{
    double t = orc_dihedral(x + 3 * hinge[0], x + 3 * hinge[1], x + 3 * hinge[2],
                            x + 3 * hinge[3], NULL, NULL);
    return bend * (t - theta0) * (t - theta0) * len0;
}

void orc_bend_grad_hess(const double* x, const uint32_t hinge[4], double theta0, double len0,
                        double bend, double dt, double g[12], double H[144])
{
    double gt[12], Ht[144];
    double t = orc_dihedral(x + 3 * hinge[0], x + 3 * hinge[1], x + 3 * hinge[2],
                            x + 3 * hinge[3], gt, Ht);
    double d = t - theta0;
    double s = bend * len0 * dt * dt;
    for(int a = 0; a < 12; a++)
    {
        g[a] = 2.0 * d * gt[a] * s;
        for(int b = 0; b < 12; b++)
            H[a * 12 + b] = 2.0 * (d * Ht[a * 12 + b] + gt[a] * gt[b]);
    }
    /* makePDSNK: unchanged when the smallest eigenvalue is >= 0 */
    double wv[12], Vv[144];
    orc_sym_eig(12, H, wv, Vv);
    if(wv[0] < 0.0)
        orc_psd_project(12, H);
    for(int k = 0; k < 144; k++)
        H[k] *= s;
}

typedef struct
{
    uint32_t a, b, opp;
    int      order;
} half_t;

static int half_cmp(const void* pa, const void* pb)
{
    const half_t* x = (const half_t*)pa;
    const half_t* y = (const half_t*)pb;
    if(x->a != y->a)
        return x->a < y->a ? -1 : 1;
    if(x->b != y->b)
        return x->b < y->b ? -1 : 1;
    return x->order - y->order;
}

int orc_bend_hinges(int ntri, const uint32_t* tris, uint32_t* hinges)
{
    half_t* h = (half_t*)malloc(sizeof(half_t) * 3 * (ntri ? ntri : 1));
    int     n = 0;
    for(int f = 0; f < ntri; f++)
        for(int k = 0; k < 3; k++)
        {
            uint32_t v0 = tris[3 * f + k], v1 = tris[3 * f + (k + 1) % 3],
                     vo = tris[3 * f + (k + 2) % 3];
            h[n].a      = v0 < v1 ? v0 : v1;
            h[n].b      = v0 < v1 ? v1 : v0;
            h[n].opp    = vo;
            h[n].order  = n;
            n++;
        }
    qsort(h, n, sizeof(half_t), half_cmp);
    int m = 0;
    for(int i = 0; i < n;)
    {
        int j = i;
        while(j < n && h[j].a == h[i].a && h[j].b == h[i].b)
            j++;
        if(j - i == 2)
        {
            if(hinges)
            {
                hinges[4 * m]     = h[i].a;
                hinges[4 * m + 1] = h[i].b;
                hinges[4 * m + 2] = h[i].opp;
                hinges[4 * m + 3] = h[i + 1].opp;
            }
            m++;
        }
        i = j;
    }
    free(h);
    return m;
}
