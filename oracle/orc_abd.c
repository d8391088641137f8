/*
 * orc_abd.c -- affine-body dynamics terms.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates StiffGIPC/abd_system:
 *   q = (p, a1, a2, a3), x = p + A xbar with A's rows a1, a2, a3
 *     (abd_jacobi_matrix.h:8-62, J = [I | xbar^T (x) rows])
 *   shape energy ||A A^T - I||_F^2 = sum (a_i.a_i - 1)^2 + 2 sum_{i<j} (a_i.a_j)^2,
 *     gradient and Hessian (details/abd_energy.inl:9-127), scaled by
 *     kappa V dt^2 and made PSD (make_pd, setup_abd_system_gradient_and_hessian.cu:180-193)
 *   dyadic mass M = sum_i m_i J_i^T J_i (_setup_abd_dyadic_mass, abd_system.cu)
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

double orc_abd_shape_energy(const double q[12])
{
    const double* a[3] = {q + 3, q + 6, q + 9};
    double        e    = 0.0;
    for(int i = 0; i < 3; i++)
        for(int j = 0; j < 3; j++)
        {
            double d = a[i][0] * a[j][0] + a[i][1] * a[j][1] + a[i][2] * a[j][2] - (i == j);
            e += d * d;
        }
    return e;
}

void orc_abd_shape_grad(const double q[12], double g[9])
{
    const double* a[3] = {q + 3, q + 6, q + 9};
    for(int i = 0; i < 3; i++)
        for(int c = 0; c < 3; c++)
        {
            double s = 0.0;
            for(int j = 0; j < 3; j++)
            {
                double d = a[i][0] * a[j][0] + a[i][1] * a[j][1] + a[i][2] * a[j][2] - (i == j);
                s += 4.0 * d * a[j][c];
            }
            g[3 * i + c] = s;
        }
}

/* d2/da_i,c da_j,d of sum_{k,l} (a_k.a_l - delta_kl)^2:
 *   i == j: 8 a_i a_i^T + 4 (a_i.a_i - 1) I + 4 sum_{k != i} a_k a_k^T
 *   i != j: 4 a_j a_i^T + 4 (a_i.a_j) I                                     */
void orc_abd_shape_hess(const double q[12], double H[81])
{
    const double* a[3] = {q + 3, q + 6, q + 9};
    memset(H, 0, sizeof(double) * 81);
    for(int i = 0; i < 3; i++)
        for(int j = 0; j < 3; j++)
            for(int c = 0; c < 3; c++)
                for(int d = 0; d < 3; d++)
                {
                    double h;
                    if(i == j)
                    {
                        double sq = a[i][0] * a[i][0] + a[i][1] * a[i][1] + a[i][2] * a[i][2];
                        h         = 8.0 * a[i][c] * a[i][d] + 4.0 * (sq - 1.0) * (c == d);
                        for(int k = 0; k < 3; k++)
                            if(k != i)
                                h += 4.0 * a[k][c] * a[k][d];
                    }
                    else
                    {
                        double dij = a[i][0] * a[j][0] + a[i][1] * a[j][1] + a[i][2] * a[j][2];
                        h          = 4.0 * a[j][c] * a[i][d] + 4.0 * dij * (c == d);
                    }
                    H[(3 * i + c) * 9 + 3 * j + d] = h;
                }
}

void orc_abd_x(const double xbar[3], const double q[12], double x[3])
{
    for(int k = 0; k < 3; k++)
        x[k] = q[k] + q[3 + 3 * k] * xbar[0] + q[4 + 3 * k] * xbar[1] + q[5 + 3 * k] * xbar[2];
}

void orc_abd_jt(const double xbar[3], const double v[3], double out[12])
{
    for(int k = 0; k < 3; k++)
    {
        out[k] = v[k];
        for(int l = 0; l < 3; l++)
            out[3 + 3 * k + l] = v[k] * xbar[l];
    }
}

/* J (3x12, row-major) of a point with body-frame coordinates xbar */
void orc_abd_jmat(const double xbar[3], double J[36])
{
    memset(J, 0, sizeof(double) * 36);
    for(int k = 0; k < 3; k++)
    {
        J[k * 12 + k] = 1.0;
        for(int l = 0; l < 3; l++)
            J[k * 12 + 3 + 3 * k + l] = xbar[l];
    }
}
