/*
 * orc_linsys.c -- BSR merge, SpMV, block-Jacobi, MAS preconditioner and PCG.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * MAS restatement (MASPreconditioner.cu, main.cu:47-75):
 *   - vertices of bank p occupy slots 16p + k in run order (setMAS_partition)
 *   - levelnum = computeNumLevels(nv)                     (:1720-1739)
 *   - level l+1 nodes = connected components of the graph induced on each
 *     16-node bank of level l, numbered bank by bank in order of each
 *     component's lowest lane                             (:157-214, :271-328,
 *                                                          :401-590)
 *   - graph = off-diagonal block pattern of A (mesh adjacency plus the
 *     contact cliques of _buildCollisionConnection_new :1128-1538)
 *   - bank matrices = (P_l^T A P_l) restricted to each bank, zero diagonal
 *     entries replaced by 1, inverted                     (:1810-2084, :628-724)
 *   - z = sum_l P_l Binv_l P_l^T r                         (:727-1024, :848-877)
 * The oracle forms every level's bank matrices by direct summation and
 * inverts with pivoted Gauss-Jordan, independent of the reference's atomic
 * scatter and warp reductions.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define BANK 16

/* ------------------------------------------------------------------ BSR */
typedef struct
{
    long long key;
    int       src;
    int       trans;
} ent_t;

static int cmp_ent(const void* a, const void* b)
{
    const ent_t* x = (const ent_t*)a;
    const ent_t* y = (const ent_t*)b;
    if(x->key != y->key)
        return (x->key > y->key) - (x->key < y->key);
    return (x->src > y->src) - (x->src < y->src);
}

int orc_bsr_from_upper_triplets(int nb, int ntrip, const int* rows, const int* cols,
                                const double* vals, orc_bsr* out)
{
    ent_t* e  = (ent_t*)malloc(sizeof(ent_t) * (size_t)(2 * ntrip + 1));
    int    ne = 0;
    for(int t = 0; t < ntrip; t++)
    {
        int r = rows[t], c = cols[t];
        if(r < 0 || c < 0 || r >= nb || c >= nb || r > c)
        {
            free(e);
            return -1;
        }
        e[ne].key   = (long long)r * nb + c;
        e[ne].src   = t;
        e[ne].trans = 0;
        ne++;
        if(r != c)
        {
            e[ne].key   = (long long)c * nb + r;
            e[ne].src   = t;
            e[ne].trans = 1;
            ne++;
        }
    }
    qsort(e, ne, sizeof(ent_t), cmp_ent);
    int nu = 0;
    for(int i = 0; i < ne; i++)
        if(i == 0 || e[i].key != e[i - 1].key)
            nu++;
    out->nb     = nb;
    out->nbody  = 0;
    out->nnzb   = nu;
    out->rowptr = (int*)calloc(nb + 1, sizeof(int));
    out->cols   = (int*)malloc(sizeof(int) * (nu ? nu : 1));
    out->vals   = (double*)calloc((size_t)9 * (nu ? nu : 1), sizeof(double));
    int u       = -1;
    for(int i = 0; i < ne; i++)
    {
        if(i == 0 || e[i].key != e[i - 1].key)
        {
            u++;
            int r        = (int)(e[i].key / nb);
            out->cols[u] = (int)(e[i].key % nb);
            out->rowptr[r + 1]++;
        }
        const double* s = vals + 9 * e[i].src;
        double*       d = out->vals + 9 * u;
        for(int a = 0; a < 3; a++)
            for(int b = 0; b < 3; b++)
                d[a * 3 + b] += e[i].trans ? s[b * 3 + a] : s[a * 3 + b];
    }
    for(int r = 0; r < nb; r++)
        out->rowptr[r + 1] += out->rowptr[r];
    free(e);
    return 0;
}

void orc_bsr_free(orc_bsr* A)
{
    free(A->rowptr);
    free(A->cols);
    free(A->vals);
    A->rowptr = NULL;
    A->cols   = NULL;
    A->vals   = NULL;
}

void orc_bsr_spmv(const orc_bsr* A, const double* x, double* y)
{
#pragma omp parallel for schedule(static)
    for(int i = 0; i < A->nb; i++)
    {
        double acc[3] = {0, 0, 0};
        for(int k = A->rowptr[i]; k < A->rowptr[i + 1]; k++)
        {
            const double* b = A->vals + 9 * k;
            const double* xj = x + 3 * A->cols[k];
            for(int a = 0; a < 3; a++)
                acc[a] += b[a * 3 + 0] * xj[0] + b[a * 3 + 1] * xj[1] + b[a * 3 + 2] * xj[2];
        }
        y[3 * i + 0] = acc[0];
        y[3 * i + 1] = acc[1];
        y[3 * i + 2] = acc[2];
    }
}

/* ------------------------------------------------------------------ MAS */
struct orc_mas
{
    int     nv, nparts, total_map, levels, total_clusters;
    int*    slot;        /* nv: level-0 slot of vertex                */
    int*    node;        /* levels x nv: local node id at level l>=1  */
    int     level_count[8];
    int     level_off[8];
    double* inv;         /* (total_clusters/16) x 48 x 48 */
    int     fp32;
};

static int compute_num_levels(int nv)
{
    int nLevel  = 1;
    int levelSz = (nv + BANK - 1) / BANK * BANK;
    while(levelSz > BANK)
    {
        levelSz /= BANK;
        nLevel++;
        levelSz = (levelSz + BANK - 1) / BANK * BANK;
    }
    nLevel = nLevel + 1;
    return nLevel > 6 ? 6 : nLevel;
}

static int uf_find(int* p, int a)
{
    while(p[a] != a)
    {
        p[a] = p[p[a]];
        a    = p[a];
    }
    return a;
}

/* cluster the nodes [0, n) of one level: components of edges (a,b) whose
 * endpoints share a bank.  Returns the number of clusters; newid[n]. */
static int cluster_level(int n, int nvalid_mask_len, const int* valid, int ne, const int* ea,
                         const int* eb, int* newid)
{
    int* p = (int*)malloc(sizeof(int) * (n ? n : 1));
    for(int i = 0; i < n; i++)
        p[i] = i;
    for(int k = 0; k < ne; k++)
    {
        int a = ea[k], b = eb[k];
        if(a < 0 || b < 0 || a == b || a / BANK != b / BANK)
            continue;
        int ra = uf_find(p, a), rb = uf_find(p, b);
        if(ra != rb)
        {
            if(ra < rb)
                p[rb] = ra;
            else
                p[ra] = rb;
        }
    }
    /* root of each component is its lowest member (union by min) */
    int cnt = 0;
    for(int i = 0; i < n; i++)
    {
        if(valid && !valid[i])
        {
            newid[i] = -1;
            continue;
        }
        int r = uf_find(p, i);
        if(r == i)
            newid[i] = cnt++;
        else
            newid[i] = newid[r];
    }
    (void)nvalid_mask_len;
    free(p);
    return cnt;
}

orc_mas* orc_mas_create(const orc_bsr* A, const int32_t* part_id, int fp32)
{
    int      nv = A->nb;
    orc_mas* m  = (orc_mas*)calloc(1, sizeof(orc_mas));
    m->nv       = nv;
    m->fp32     = fp32;
    m->slot     = (int*)malloc(sizeof(int) * nv);
    /* slots (setMAS_partition, main.cu:47-75) */
    int part = -1, idx = 0, maxpart = -1;
    for(int v = 0; v < nv; v++)
    {
        int pv = part_id ? part_id[v] : v / BANK;
        if(pv != part)
        {
            part = pv;
            idx  = 0;
        }
        m->slot[v] = BANK * pv + idx;
        idx++;
        if(pv > maxpart)
            maxpart = pv;
    }
    m->nparts    = maxpart + 1;
    m->total_map = BANK * m->nparts;
    m->levels    = compute_num_levels(nv);
    int L        = m->levels;

    /* edges of the fine graph (each unordered pair once) */
    int  ne = 0;
    int* ea = (int*)malloc(sizeof(int) * (A->nnzb + 1));
    int* eb = (int*)malloc(sizeof(int) * (A->nnzb + 1));
    for(int i = 0; i < nv; i++)
        for(int k = A->rowptr[i]; k < A->rowptr[i + 1]; k++)
        {
            int j = A->cols[k];
            if(j > i)
            {
                ea[ne] = i;
                eb[ne] = j;
                ne++;
            }
        }

    m->node = (int*)malloc(sizeof(int) * (size_t)L * nv);
    int* ta = (int*)malloc(sizeof(int) * (ne + 1));
    int* tb = (int*)malloc(sizeof(int) * (ne + 1));

    /* level 0 -> 1 on slots */
    int* valid = (int*)calloc(m->total_map, sizeof(int));
    int* nid   = (int*)malloc(sizeof(int) * (m->total_map + 1));
    for(int v = 0; v < nv; v++)
        valid[m->slot[v]] = 1;
    for(int k = 0; k < ne; k++)
    {
        ta[k] = m->slot[ea[k]];
        tb[k] = m->slot[eb[k]];
    }
    int n1                = cluster_level(m->total_map, 0, valid, ne, ta, tb, nid);
    m->level_count[0]     = nv;
    m->level_off[0]       = 0;
    m->level_count[1]     = n1;
    m->level_off[1]       = (m->total_map + BANK - 1) / BANK * BANK;
    for(int v = 0; v < nv; v++)
        m->node[1 * nv + v] = nid[m->slot[v]];
    /* levels 1 .. L-1 -> next */
    for(int l = 1; l < L; l++)
    {
        int n = m->level_count[l];
        int* nid2 = (int*)malloc(sizeof(int) * (n + 1));
        for(int k = 0; k < ne; k++)
        {
            ta[k] = m->node[l * nv + ea[k]];
            tb[k] = m->node[l * nv + eb[k]];
        }
        int nn                = cluster_level(n, 0, NULL, ne, ta, tb, nid2);
        m->level_count[l + 1] = nn;
        m->level_off[l + 1]   = m->level_off[l] + (n + BANK - 1) / BANK * BANK;
        if(l + 1 < L)
            for(int v = 0; v < nv; v++)
                m->node[(l + 1) * nv + v] = nid2[m->node[l * nv + v]];
        free(nid2);
    }
    m->total_clusters = m->level_off[L];

    /* bank matrices */
    int     nbanks = m->total_clusters / BANK;
    double* B      = (double*)calloc((size_t)nbanks * 48 * 48, sizeof(double));
    for(int i = 0; i < nv; i++)
        for(int k = A->rowptr[i]; k < A->rowptr[i + 1]; k++)
        {
            int           j   = A->cols[k];
            const double* blk = A->vals + 9 * k;
            for(int l = 0; l < L; l++)
            {
                int gi, gj;
                if(l == 0)
                {
                    gi = m->slot[i];
                    gj = m->slot[j];
                }
                else
                {
                    gi = m->level_off[l] + m->node[l * nv + i];
                    gj = m->level_off[l] + m->node[l * nv + j];
                }
                if(gi / BANK != gj / BANK)
                    continue;
                double* M  = B + (size_t)(gi / BANK) * 48 * 48;
                int     li = gi % BANK, lj = gj % BANK;
                for(int a = 0; a < 3; a++)
                    for(int b = 0; b < 3; b++)
                        M[(3 * li + a) * 48 + 3 * lj + b] += blk[a * 3 + b];
            }
        }
#pragma omp parallel for schedule(dynamic, 16)
    for(int bk = 0; bk < nbanks; bk++)
    {
        double* M = B + (size_t)bk * 48 * 48;
        for(int d = 0; d < 48; d++)
            if(M[d * 48 + d] == 0.0)
                M[d * 48 + d] = 1.0;
        orc_inverse_gj(48, M);
        /* the diagonal 3x3 blocks take their upper entries below the diagonal
         * too (MASPreconditioner.cu:704-708) */
        for(int I = 0; I < BANK; I++)
        {
            double* D = M + (size_t)(3 * I) * 48 + 3 * I;
            D[1 * 48 + 0] = D[0 * 48 + 1];
            D[2 * 48 + 1] = D[1 * 48 + 2];
            D[2 * 48 + 0] = D[0 * 48 + 2];
        }
        if(fp32)
            for(int e = 0; e < 48 * 48; e++)
                M[e] = (double)(float)M[e];
        /* only the upper 3x3 blocks are kept (MasMatrixSymf, eigen_data.h:125-128,
         * stored by __inverse6_P96x96, MASPreconditioner.cu:628-724); the apply
         * uses their transposes below the diagonal */
        for(int I = 1; I < BANK; I++)
            for(int J = 0; J < I; J++)
                for(int a = 0; a < 3; a++)
                    for(int b = 0; b < 3; b++)
                        M[(3 * I + a) * 48 + 3 * J + b] = M[(3 * J + b) * 48 + 3 * I + a];
    }
    m->inv = B;
    free(ea);
    free(eb);
    free(ta);
    free(tb);
    free(valid);
    free(nid);
    return m;
}

void orc_mas_free(orc_mas* m)
{
    if(!m)
        return;
    free(m->slot);
    free(m->node);
    free(m->inv);
    free(m);
}

int orc_mas_levels(const orc_mas* m) { return m->levels; }
int orc_mas_total_clusters(const orc_mas* m) { return m->total_clusters; }
void orc_mas_level_sizes(const orc_mas* m, int32_t* s)
{
    for(int l = 0; l <= m->levels; l++)
    {
        s[2 * l]     = m->level_count[l];
        s[2 * l + 1] = m->level_off[l];
    }
}
void orc_mas_chain(const orc_mas* m, int32_t* chain)
{
    int L = m->levels;
    for(int v = 0; v < m->nv; v++)
        for(int l = 1; l < L; l++)
            chain[v * (L - 1) + l - 1] = m->level_off[l] + m->node[l * m->nv + v];
}

void orc_mas_apply(const orc_mas* m, const double* r, double* z)
{
    int     nv = m->nv, L = m->levels, tc = m->total_clusters;
    double* R  = (double*)calloc((size_t)3 * tc, sizeof(double));
    double* Z  = (double*)calloc((size_t)3 * tc, sizeof(double));
    for(int v = 0; v < nv; v++)
        for(int l = 0; l < L; l++)
        {
            int g = (l == 0) ? m->slot[v] : m->level_off[l] + m->node[l * nv + v];
            for(int a = 0; a < 3; a++)
                R[3 * g + a] += r[3 * v + a];
        }
    if(m->fp32)
        for(int i = 0; i < 3 * tc; i++)
            R[i] = (double)(float)R[i];
#pragma omp parallel for schedule(static)
    for(int bk = 0; bk < tc / BANK; bk++)
    {
        const double* M = m->inv + (size_t)bk * 48 * 48;
        for(int i = 0; i < 48; i++)
        {
            double s = 0.0;
            for(int j = 0; j < 48; j++)
                s += M[i * 48 + j] * R[48 * bk + j];
            Z[48 * bk + i] = s;
        }
    }
#pragma omp parallel for schedule(static)
    for(int v = 0; v < nv; v++)
    {
        double acc[3] = {0, 0, 0};
        for(int l = 0; l < L; l++)
        {
            int g = (l == 0) ? m->slot[v] : m->level_off[l] + m->node[l * nv + v];
            for(int a = 0; a < 3; a++)
                acc[a] += Z[3 * g + a];
        }
        for(int a = 0; a < 3; a++)
            z[3 * v + a] = acc[a];
    }
    free(R);
    free(Z);
}

/* ------------------------------------------------------------------ PCG */
/* DiagPreconditioner (diag_preconditioner.cu:9-66) on the nodes before the
 * affine bodies, then ABDPreconditioner (abd_preconditioner.cu:15-38,
 * setup_abd_system_gradient_and_hessian.cu:603-646) on each body's 4 nodes:
 * the inverse of its 12x12 diagonal block (a singular block: identity),
 * applied row by row in index order */
static void block_jacobi(const orc_bsr* A, const double* r, double* z)
{
    int nb3 = A->nb - 4 * A->nbody;
    for(int b = 0; b < A->nbody; b++)
    {
        int    n0 = nb3 + 4 * b;
        double M[144];
        for(int m = 0; m < 4; m++)
            for(int n = 0; n < 4; n++)
            {
                const double* blk = NULL;
                for(int k = A->rowptr[n0 + m]; k < A->rowptr[n0 + m + 1]; k++)
                    if(A->cols[k] == n0 + n)
                        blk = A->vals + 9 * k;
                for(int i = 0; i < 3; i++)
                    for(int j = 0; j < 3; j++)
                        M[(3 * m + i) * 12 + 3 * n + j] = blk ? blk[3 * i + j] : 0.0;
            }
        if(orc_inverse(12, M) != 0)
            for(int i = 0; i < 144; i++)
                M[i] = (i % 13 == 0) ? 1.0 : 0.0;
        for(int i = 0; i < 12; i++)
        {
            double zi = 0.0;
            for(int k = 0; k < 12; k++)
                zi += M[i * 12 + k] * r[3 * n0 + k];
            z[3 * n0 + i] = zi;
        }
    }
    for(int i = 0; i < nb3; i++)
    {
        double D[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        for(int k = A->rowptr[i]; k < A->rowptr[i + 1]; k++)
            if(A->cols[k] == i)
                memcpy(D, A->vals + 9 * k, sizeof(D));
        orc_inverse(3, D);
        for(int a = 0; a < 3; a++)
            z[3 * i + a] = D[a * 3 + 0] * r[3 * i + 0] + D[a * 3 + 1] * r[3 * i + 1]
                           + D[a * 3 + 2] * r[3 * i + 2];
    }
}

void orc_block_jacobi(const orc_bsr* A, const double* r, double* z) { block_jacobi(A, r, z); }

static void precond(const orc_bsr* A, const orc_mas* mas, int pc, const double* r, double* z)
{
    if(pc == 1 && mas)
        orc_mas_apply(mas, r, z);
    else if(pc == 0)
        block_jacobi(A, r, z);
    else
        memcpy(z, r, sizeof(double) * 3 * A->nb);
}

static double dot(int n, const double* a, const double* b)
{
    double s = 0.0;
    for(int i = 0; i < n; i++)
        s += a[i] * b[i];
    return s;
}

int orc_pcg(const orc_bsr* A, const orc_mas* mas, int pc, const double* b, double* x,
            double rel_tol, double abs_tol, int max_iter, int use_rz, int check_every,
            double* rz0_out, double* rz_out)
{
    int     n  = 3 * A->nb;
    double* r  = (double*)malloc(sizeof(double) * n);
    double* z  = (double*)malloc(sizeof(double) * n);
    double* p  = (double*)malloc(sizeof(double) * n);
    double* Ap = (double*)malloc(sizeof(double) * n);
    if(max_iter <= 0)
        max_iter = n;
    if(check_every <= 0)
        check_every = 10;
    memset(x, 0, sizeof(double) * n);
    memcpy(r, b, sizeof(double) * n);
    precond(A, mas, pc, r, z);
    double rz = dot(n, r, z);
    memcpy(p, z, sizeof(double) * n);
    double rz0 = rz, rr0 = use_rz ? 0.0 : dot(n, r, r);
    int    k   = 0;
    if(rz0 == 0.0)
    {
        if(rz0_out)
            *rz0_out = 0.0;
        if(rz_out)
            *rz_out = 0.0;
        free(r);
        free(z);
        free(p);
        free(Ap);
        return 0;
    }
    for(k = 1; k < max_iter; k++)
    {
        orc_bsr_spmv(A, p, Ap);
        double alpha = rz / dot(n, p, Ap);
#pragma omp parallel for schedule(static)
        for(int i = 0; i < n; i++)
        {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
        }
        if(k % check_every == 0)
        {
            orc_bsr_spmv(A, x, Ap);
            for(int i = 0; i < n; i++)
                r[i] = b[i] - Ap[i];
        }
        precond(A, mas, pc, r, z);
        double rzn = dot(n, r, z);
        if(k % check_every == 0)
        {
            double rel2 = rel_tol * rel_tol, abs2 = abs_tol * abs_tol;
            if(use_rz)
            {
                if(rzn <= rel2 * rz0 || rzn <= abs2)
                {
                    rz = rzn;
                    break;
                }
            }
            else
            {
                double rr = dot(n, r, r);
                if(rr <= rel2 * rr0 || rr <= abs2)
                {
                    rz = rzn;
                    break;
                }
            }
        }
        double beta = rzn / rz;
#pragma omp parallel for schedule(static)
        for(int i = 0; i < n; i++)
            p[i] = z[i] + beta * p[i];
        rz = rzn;
    }
    if(rz0_out)
        *rz0_out = rz0;
    if(rz_out)
        *rz_out = rz;
    free(r);
    free(z);
    free(p);
    free(Ap);
    return k;
}
