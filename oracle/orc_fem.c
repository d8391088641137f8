/*
 * orc_fem.c -- FEM (SNK1 neo-Hookean), kinetic and ground-barrier terms.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * SNK1 (the model the reference compiles, CMakeLists.txt:58 USE_SNK1):
 *   psi(F) = mu/2 (|F|^2 - 3) - mu ln J + lam/2 ln^2 J
 *   (femEnergy.cu:1070-1088 energy, :443-486 PK1, :1706-1786 Hessian).
 * The oracle builds the full analytic 9x9 Hessian in vec(F) space
 *   H = mu I + lam G(x)G + (mu - lam lnJ) T,  G = F^-T,
 *   T[(ij),(kl)] = G_il G_kj
 * and projects it with a generic Jacobi eigen-solver, which is independent of
 * the reference's (and the HIP path's) closed-form twist/flip eigenbasis.
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

static double det3(const double* F)
{
    return F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6])
           + F[2] * (F[3] * F[7] - F[4] * F[6]);
}

/* G = F^{-T} */
static void inv_transpose(const double* F, double* G)
{
    double J = det3(F);
    /* cofactor matrix C_ij; F^{-T} = C / J */
    G[0] = (F[4] * F[8] - F[5] * F[7]) / J;
    G[1] = -(F[3] * F[8] - F[5] * F[6]) / J;
    G[2] = (F[3] * F[7] - F[4] * F[6]) / J;
    G[3] = -(F[1] * F[8] - F[2] * F[7]) / J;
    G[4] = (F[0] * F[8] - F[2] * F[6]) / J;
    G[5] = -(F[0] * F[7] - F[1] * F[6]) / J;
    G[6] = (F[1] * F[5] - F[2] * F[4]) / J;
    G[7] = -(F[0] * F[5] - F[2] * F[3]) / J;
    G[8] = (F[0] * F[4] - F[1] * F[3]) / J;
}

double orc_snk1_psi(const double F[9], double mu, double lam)
{
    double I1 = 0.0;
    for(int i = 0; i < 9; i++)
        I1 += F[i] * F[i];
    double J  = det3(F);
    double lj = log(J);
    return 0.5 * mu * (I1 - 3.0) - mu * lj + 0.5 * lam * lj * lj;
}

void orc_snk1_pk1(const double F[9], double mu, double lam, double P[9])
{
    double G[9];
    inv_transpose(F, G);
    double c = lam * log(det3(F)) - mu;
    for(int i = 0; i < 9; i++)
        P[i] = mu * F[i] + c * G[i];
}

void orc_snk1_hess(const double F[9], double mu, double lam, int project, double H[81])
{
    double G[9];
    inv_transpose(F, G);
    double c = lam * log(det3(F)) - mu;
    /* vec index a = i + 3 j for F_ij (column-major vec, as the reference) */
    for(int j = 0; j < 3; j++)
        for(int i = 0; i < 3; i++)
            for(int l = 0; l < 3; l++)
                for(int k = 0; k < 3; k++)
                {
                    int a = i + 3 * j, b = k + 3 * l;
                    double h = lam * G[i * 3 + j] * G[k * 3 + l]
                               - c * G[i * 3 + l] * G[k * 3 + j];
                    if(a == b)
                        h += mu;
                    H[a * 9 + b] = h;
                }
    if(project)
        orc_psd_project(9, H);
}

static void tet_F(const double* x, const uint32_t t[4], const double dminv[9], double F[9])
{
    double Ds[9];
    for(int i = 0; i < 3; i++)
        for(int k = 0; k < 3; k++)
            Ds[i * 3 + k] = x[3 * t[k + 1] + i] - x[3 * t[0] + i];
    for(int i = 0; i < 3; i++)
        for(int j = 0; j < 3; j++)
        {
            double s = 0.0;
            for(int k = 0; k < 3; k++)
                s += Ds[i * 3 + k] * dminv[k * 3 + j];
            F[i * 3 + j] = s;
        }
}

/* B[a][j] = dF_ij / dx_{a,i}  (femEnergy.cu:1438-1486 PFPX, compact form) */
static void tet_B(const double dminv[9], double B[12])
{
    for(int j = 0; j < 3; j++)
    {
        B[0 * 3 + j] = -(dminv[0 * 3 + j] + dminv[1 * 3 + j] + dminv[2 * 3 + j]);
        for(int k = 0; k < 3; k++)
            B[(k + 1) * 3 + j] = dminv[k * 3 + j];
    }
}

void orc_fem_precompute(int nv, const double* xr, int nt, const uint32_t* tets,
                        const double* E, const double* nu, const double* rho,
                        double* dminv, double* vol, double* mu, double* lam, double* mass)
{
    for(int v = 0; v < nv; v++)
        mass[v] = 0.0;
    for(int e = 0; e < nt; e++)
    {
        const uint32_t* t = tets + 4 * e;
        double Dm[9];
        for(int i = 0; i < 3; i++)
            for(int k = 0; k < 3; k++)
                Dm[i * 3 + k] = xr[3 * t[k + 1] + i] - xr[3 * t[0] + i];
        double V = fabs(det3(Dm)) / 6.0;
        orc_inverse(3, Dm);
        memcpy(dminv + 9 * e, Dm, sizeof(double) * 9);
        vol[e] = V;
        double m = V * rho[e] / 4.0;
        for(int k = 0; k < 4; k++)
            mass[t[k]] += m;
        mu[e]  = E[e] / (2.0 * (1.0 + nu[e]));
        lam[e] = E[e] * nu[e] / ((1.0 + nu[e]) * (1.0 - 2.0 * nu[e]));
    }
}

double orc_tet_energy(const double* x, const uint32_t t[4], const double dminv[9],
                      double vol, double mu, double lam)
{
    double F[9];
    tet_F(x, t, dminv, F);
    return vol * orc_snk1_psi(F, mu, lam);
}

void orc_tet_grad_hess(const double* x, const uint32_t t[4], const double dminv[9],
                       double vol, double mu, double lam, double dt, double g[12],
                       double H[144])
{
    double F[9], P[9], Hq[81], B[12];
    tet_F(x, t, dminv, F);
    orc_snk1_pk1(F, mu, lam, P);
    orc_snk1_hess(F, mu, lam, 1, Hq);
    tet_B(dminv, B);
    double s = dt * dt * vol;
    for(int a = 0; a < 4; a++)
        for(int i = 0; i < 3; i++)
        {
            double acc = 0.0;
            for(int j = 0; j < 3; j++)
                acc += P[i * 3 + j] * B[a * 3 + j];
            g[3 * a + i] = s * acc;
        }
    for(int a = 0; a < 4; a++)
        for(int i = 0; i < 3; i++)
            for(int b = 0; b < 4; b++)
                for(int m = 0; m < 3; m++)
                {
                    double acc = 0.0;
                    for(int j = 0; j < 3; j++)
                        for(int l = 0; l < 3; l++)
                            acc += B[a * 3 + j] * B[b * 3 + l] * Hq[(i + 3 * j) * 9 + (m + 3 * l)];
                    H[(3 * a + i) * 12 + 3 * b + m] = s * acc;
                }
}

double orc_ground_energy(double dist2, double dhat2)
{
    double t = dist2 - dhat2;
    return -t * t * log(dist2 / dhat2);
}

void orc_ground_grad_hess(const double x[3], const double n[3], double off, double dhat2,
                          double kappa, double g[3], double H[9])
{
    double d  = n[0] * x[0] + n[1] * x[1] + n[2] * x[2] - off;
    double d2 = d * d;
    double t  = d2 - dhat2;
    /* b(s) = -(s-dhat2)^2 ln(s/dhat2), s = d^2 */
    double bp  = -2.0 * t * log(d2 / dhat2) - t * t / d2;
    double bpp = -2.0 * log(d2 / dhat2) - 4.0 * t / d2 + t * t / (d2 * d2);
    for(int i = 0; i < 3; i++)
        g[i] = kappa * bp * 2.0 * d * n[i];
    double p = kappa * (4.0 * bpp * d2 + 2.0 * bp);
    for(int i = 0; i < 3; i++)
        for(int j = 0; j < 3; j++)
            H[i * 3 + j] = p * n[i] * n[j];
}
