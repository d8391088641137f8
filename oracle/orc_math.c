/*
 * orc_math.c -- small dense linear algebra for the oracle.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference relies on Eigen::SelfAdjointEigenSolver for PSD projection
 * (GIPC.cu:33-80, femEnergy.cu:67-85).  Eigen is not vendored; the oracle uses
 * the cyclic Jacobi method (Golub & Van Loan, Matrix Computations, 4th ed.,
 * Alg. 8.5.3), which converges to full FP64 accuracy for symmetric matrices.
 */
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include "oracle.h"

void orc_sym_eig(int n, const double* Ain, double* w, double* V)
{
    double A[144];
    for(int i = 0; i < n * n; i++)
        A[i] = Ain[i];
    for(int i = 0; i < n; i++)
        for(int j = 0; j < n; j++)
            V[i * n + j] = (i == j) ? 1.0 : 0.0;

    for(int sweep = 0; sweep < 100; sweep++)
    {
        double off = 0.0, tot = 0.0;
        for(int i = 0; i < n; i++)
            for(int j = 0; j < n; j++)
            {
                tot += A[i * n + j] * A[i * n + j];
                if(i != j)
                    off += A[i * n + j] * A[i * n + j];
            }
        if(off <= 1e-30 * tot || off == 0.0)
            break;
        for(int p = 0; p < n - 1; p++)
            for(int q = p + 1; q < n; q++)
            {
                double apq = A[p * n + q];
                if(apq == 0.0)
                    continue;
                double app = A[p * n + p], aqq = A[q * n + q];
                double tau = (aqq - app) / (2.0 * apq);
                double t   = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                double c   = 1.0 / sqrt(1.0 + t * t);
                double s   = t * c;
                for(int k = 0; k < n; k++)
                {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for(int k = 0; k < n; k++)
                {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for(int k = 0; k < n; k++)
                {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    for(int i = 0; i < n; i++)
        w[i] = A[i * n + i];
    /* sort ascending (selection sort keeps columns paired) */
    for(int i = 0; i < n - 1; i++)
    {
        int m = i;
        for(int j = i + 1; j < n; j++)
            if(w[j] < w[m])
                m = j;
        if(m != i)
        {
            double t = w[i];
            w[i]     = w[m];
            w[m]     = t;
            for(int k = 0; k < n; k++)
            {
                double tv    = V[k * n + i];
                V[k * n + i] = V[k * n + m];
                V[k * n + m] = tv;
            }
        }
    }
}

void orc_psd_project(int n, double* A)
{
    double w[12], V[144];
    orc_sym_eig(n, A, w, V);
    if(w[0] >= 0.0)
        return;
    for(int i = 0; i < n; i++)
        for(int j = 0; j < n; j++)
        {
            double s = 0.0;
            for(int k = 0; k < n; k++)
                if(w[k] > 0.0)
                    s += V[i * n + k] * w[k] * V[j * n + k];
            A[i * n + j] = s;
        }
}

int orc_inverse(int n, double* A)
{
    int*    piv = (int*)malloc(sizeof(int) * n);
    double* B   = (double*)malloc(sizeof(double) * n * n);
    for(int i = 0; i < n * n; i++)
        B[i] = 0.0;
    for(int i = 0; i < n; i++)
        B[i * n + i] = 1.0;
    int ok = 0;
    for(int c = 0; c < n; c++)
    {
        int    p  = c;
        double mx = fabs(A[c * n + c]);
        for(int r = c + 1; r < n; r++)
            if(fabs(A[r * n + c]) > mx)
            {
                mx = fabs(A[r * n + c]);
                p  = r;
            }
        if(mx == 0.0)
        {
            ok = -1;
            break;
        }
        if(p != c)
            for(int k = 0; k < n; k++)
            {
                double t     = A[c * n + k];
                A[c * n + k] = A[p * n + k];
                A[p * n + k] = t;
                t            = B[c * n + k];
                B[c * n + k] = B[p * n + k];
                B[p * n + k] = t;
            }
        double inv = 1.0 / A[c * n + c];
        for(int k = 0; k < n; k++)
        {
            A[c * n + k] *= inv;
            B[c * n + k] *= inv;
        }
        for(int r = 0; r < n; r++)
        {
            if(r == c)
                continue;
            double f = A[r * n + c];
            if(f == 0.0)
                continue;
            for(int k = 0; k < n; k++)
            {
                A[r * n + k] -= f * A[c * n + k];
                B[r * n + k] -= f * B[c * n + k];
            }
        }
    }
    if(ok == 0)
        memcpy(A, B, sizeof(double) * n * n);
    free(piv);
    free(B);
    return ok;
}

/* In-place Gauss-Jordan inverse without pivoting in the reference's operation
 * order (__inverse6_P96x96, MASPreconditioner.cu:628-724): for each pivot j,
 * rt = M[j][j] and column j are saved, column j is replaced by e_j, row j is
 * divided by rt, then every row k != j gets rate * row j added with
 * rate = -column_j[k] (one rounded product and one rounded sum per entry). */
void orc_inverse_gj(int n, double* M)
{
    double* colm = (double*)malloc(sizeof(double) * n);
    for(int j = 0; j < n; j++)
    {
        double rt = M[j * n + j];
        for(int i = 0; i < n; i++)
            colm[i] = M[i * n + j];
        for(int i = 0; i < n; i++)
            M[i * n + j] = (i == j) ? 1.0 : 0.0;
        for(int i = 0; i < n; i++)
            M[j * n + i] /= rt;
        for(int k = 0; k < n; k++)
        {
            if(k == j)
                continue;
            double rate = -colm[k];
            for(int i = 0; i < n; i++)
            {
                double t = rate * M[j * n + i];
                M[k * n + i] += t;
            }
        }
    }
    free(colm);
}

/* Host threads of the parallel loops.  Parallel loops only compute per-index
 * results (element gradients/Hessians/energies, SpMV rows, bank inverses and
 * products); every sum over indices is then taken sequentially in index
 * order, so results are identical for any thread count. */
#include <omp.h>
void orc_set_threads(int n)
{
    omp_set_num_threads(n > 0 ? n : 1);
}
int orc_get_threads(void)
{
    return omp_get_max_threads();
}
