// context.hip -- the Newton driver (host C++ around the device kernels).
//
// Mirrors StiffGIPC/GIPC.cu:
//   set-up           initFEM (main.cu:310-356), GIPC::init (:8534-8581)
//   assembly         computeGradientAndHessian (:10012-10228)
//   one iteration    solve_subIP loop body (:10746-10862)
//   line search      lineSearch (:10573-10670)
//   kappa            suggestKappa / upperBoundKappa / initKappa (:9746-9831);
//                    postLineSearch's doubling never fires in the reference
//                    (h_close_*Num are never read back, GIPC.cuh:69-70) and is
//                    therefore not reproduced
//   time step        IPC_Solver (:10940-11131) with lagged friction
// All state stays on the device; the host reads back only the scalars that
// steer control flow (contact counts at capacity checks, step bounds, line
// search energies, PCG convergence tests).
#include <algorithm>
#include <array>
#include <cmath>

#include "cloth.h"
#include "context.h"
#include "mm_io.h"

namespace sgipc
{

// Gauss-Jordan inverse with partial pivoting (row-major n x n); false when
// singular
static bool invert_dense(int n, const double* A, double* Ai)
{
    std::vector<double> a(A, A + n * n);
    for(int i = 0; i < n * n; i++)
        Ai[i] = (i % (n + 1) == 0) ? 1.0 : 0.0;
    for(int c = 0; c < n; c++)
    {
        int p = c;
        for(int r = c + 1; r < n; r++)
            if(std::fabs(a[r * n + c]) > std::fabs(a[p * n + c]))
                p = r;
        if(a[p * n + c] == 0.0)
            return false;
        for(int k = 0; k < n; k++)
        {
            std::swap(a[c * n + k], a[p * n + k]);
            std::swap(Ai[c * n + k], Ai[p * n + k]);
        }
        double inv = 1.0 / a[c * n + c];
        for(int k = 0; k < n; k++)
        {
            a[c * n + k] *= inv;
            Ai[c * n + k] *= inv;
        }
        for(int r = 0; r < n; r++)
            if(r != c && a[r * n + c] != 0.0)
            {
                double f = a[r * n + c];
                for(int k = 0; k < n; k++)
                {
                    a[r * n + k] -= f * a[c * n + k];
                    Ai[r * n + k] -= f * Ai[c * n + k];
                }
            }
    }
    return true;
}

static double now_ms(hipEvent_t a, hipEvent_t b)
{
    float ms = 0;
    SG_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
}

Context::Context(const sgipc_scene_desc& d)
{
    desc_ = d;
    nv_   = d.n_verts;
    nt_   = d.n_tets;
    nf_   = d.n_faces;
    ne_   = d.n_edges;
    ns_   = d.n_surf_verts;
    SG_CHECK(nv_ > 0 && d.x, SGIPC_EINVAL, "scene has no vertices");
    SG_CHECK(nt_ == 0 || (d.tets && d.tet_E && d.tet_nu && d.tet_rho), SGIPC_EINVAL,
             "tet arrays missing");
    SG_CHECK(nf_ == 0 || d.faces, SGIPC_EINVAL, "faces missing");
    SG_CHECK(ne_ == 0 || d.edges, SGIPC_EINVAL, "edges missing");
    SG_CHECK(ns_ == 0 || d.surf_verts, SGIPC_EINVAL, "surface vertices missing");
    SG_CHECK(d.dt > 0 && d.dhat > 0, SGIPC_EINVAL, "dt and dhat must be positive");
    SG_CHECK(d.friction_mu >= 0.0 && d.ground_friction_mu >= 0.0, SGIPC_EINVAL,
             "friction coefficients must be non-negative");
    for(int e = 0; e < 4 * nt_; e++)
        SG_CHECK(d.tets[e] < (uint32_t)nv_, SGIPC_EINVAL, "tet vertex index out of range");
    for(int e = 0; e < 3 * nf_; e++)
        SG_CHECK(d.faces[e] < (uint32_t)nv_, SGIPC_EINVAL, "face vertex index out of range");
    for(int e = 0; e < 2 * ne_; e++)
        SG_CHECK(d.edges[e] < (uint32_t)nv_, SGIPC_EINVAL, "edge vertex index out of range");
    for(int e = 0; e < ns_; e++)
        SG_CHECK(d.surf_verts[e] < (uint32_t)nv_, SGIPC_EINVAL, "surface vertex out of range");
    SG_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    own_s_ = s_;
    dHat_  = d.dhat * d.dhat;
    kappa_ = d.kappa > 0 ? d.kappa : 0.0;

    const double*    xr = d.x_rest ? d.x_rest : d.x;
    std::vector<int> bt(nv_, 0), body(nv_, -1);
    if(d.btype)
        bt.assign(d.btype, d.btype + nv_);
    if(d.body_id)
        body.assign(d.body_id, d.body_id + nv_);

    // ---- FEM constants (initFEM): Dm^-1, |V|, Lame, lumped mass
    std::vector<double> mass(nv_, 0.0), dminv((size_t)9 * nt_), vol(nt_), mu(nt_), lam(nt_);
    for(int e = 0; e < nt_; e++)
    {
        const uint32_t* t = d.tets + 4 * (size_t)e;
        double          D[9];
        for(int i = 0; i < 3; i++)
            for(int k = 0; k < 3; k++)
                D[i * 3 + k] = xr[3 * t[k + 1] + i] - xr[3 * t[0] + i];
        double c00 = D[4] * D[8] - D[5] * D[7], c01 = D[5] * D[6] - D[3] * D[8],
               c02 = D[3] * D[7] - D[4] * D[6];
        double det = D[0] * c00 + D[1] * c01 + D[2] * c02;
        SG_CHECK(det != 0.0, SGIPC_EINVAL, "degenerate rest tetrahedron");
        double inv[9] = {c00 / det,
                         (D[2] * D[7] - D[1] * D[8]) / det,
                         (D[1] * D[5] - D[2] * D[4]) / det,
                         c01 / det,
                         (D[0] * D[8] - D[2] * D[6]) / det,
                         (D[2] * D[3] - D[0] * D[5]) / det,
                         c02 / det,
                         (D[1] * D[6] - D[0] * D[7]) / det,
                         (D[0] * D[4] - D[1] * D[3]) / det};
        for(int k = 0; k < 9; k++)
            dminv[(size_t)k * nt_ + e] = inv[k];  // SoA
        double V = std::fabs(det) / 6.0;
        vol[e]   = V;
        for(int k = 0; k < 4; k++)
            mass[t[k]] += V * d.tet_rho[e] / 4;
        double E = d.tet_E[e], nu = d.tet_nu[e];
        mu[e]    = E / (2 * (1 + nu));
        lam[e]   = E * nu / ((1 + nu) * (1 - 2 * nu));
    }
    // ---- affine bodies: vertices [nfem, nv) grouped by body; their tets
    // only supply mass and volume; q = (c, I) with c the mass centre, xbar =
    // x - c, dyadic mass M = sum m J^T J (abd_system.cu _setup_abd_state,
    // _setup_J, _setup_abd_dyadic_mass), q' and affine gravity G fitted by
    // M^-1 sum m J^T f (_setup_abd_gravity)
    std::vector<double>   hx(d.x, d.x + 3 * (size_t)nv_), hv(3 * (size_t)nv_, 0.0);
    if(d.v)
        hv.assign(d.v, d.v + 3 * (size_t)nv_);
    std::vector<int> abd_of(nv_, -1);
    nabd_ = (d.abd_body && d.n_abd > 0) ? d.n_abd : 0;
    if(nabd_)
        abd_of.assign(d.abd_body, d.abd_body + nv_);
    nfem_ = nv_;
    for(int v = 0; v < nv_; v++)
        if(abd_of[v] >= 0)
        {
            nfem_ = v;
            break;
        }
    for(int v = nfem_; v < nv_; v++)
        SG_CHECK(abd_of[v] >= 0 && abd_of[v] < nabd_
                     && (v == nfem_ ? abd_of[v] == 0 : (abd_of[v] - abd_of[v - 1] == 0
                                                       || abd_of[v] - abd_of[v - 1] == 1)),
                 SGIPC_EINVAL, "affine-body vertices must follow all others, grouped by body");
    SG_CHECK(nabd_ == 0 || abd_of[nv_ - 1] == nabd_ - 1, SGIPC_EINVAL,
             "affine body without vertices");
    nn_ = nfem_ + 4 * nabd_;
    std::vector<uint32_t> etets;
    std::vector<double>   bvol(nabd_, 0.0);
    {
        std::vector<double> d2, v2, m2, l2;
        for(int e = 0; e < nt_; e++)
        {
            const uint32_t* t = d.tets + 4 * (size_t)e;
            int na = 0;
            for(int k = 0; k < 4; k++)
                na += abd_of[t[k]] >= 0;
            SG_CHECK(na == 0 || (na == 4 && abd_of[t[1]] == abd_of[t[0]] && abd_of[t[2]] == abd_of[t[0]]
                                 && abd_of[t[3]] == abd_of[t[0]]),
                     SGIPC_EINVAL, "a tet mixes affine-body and FEM vertices");
            if(na)
            {
                bvol[abd_of[t[0]]] += vol[e];
                continue;
            }
            etets.insert(etets.end(), t, t + 4);
            v2.push_back(vol[e]);
            m2.push_back(mu[e]);
            l2.push_back(lam[e]);
        }
        int ne2 = (int)v2.size();
        d2.resize((size_t)9 * ne2);
        for(int e = 0, j = 0; e < nt_; e++)
        {
            const uint32_t* t = d.tets + 4 * (size_t)e;
            if(abd_of[t[0]] >= 0)
                continue;
            for(int k = 0; k < 9; k++)
                d2[(size_t)k * ne2 + j] = dminv[(size_t)k * nt_ + e];
            j++;
        }
        dminv.swap(d2);
        vol.swap(v2);
        mu.swap(m2);
        lam.swap(l2);
        nt_ = ne2;
    }
    std::vector<int>    bstart(nabd_ + 1, nfem_), bfix(nabd_, 0);
    std::vector<double> xbar(3 * (size_t)nv_, 0.0), bM(144 * (size_t)nabd_, 0.0),
        bq(12 * (size_t)nabd_, 0.0), bqv(12 * (size_t)nabd_, 0.0), bG(12 * (size_t)nabd_, 0.0);
    h_bMinv_.assign(144 * (size_t)nabd_, 0.0);
    for(int v = nfem_; v < nv_; v++)
        bstart[abd_of[v] + 1] = v + 1;
    // motor bodies (btype 2 on every vertex, BodyBoundaryType::Motor) move
    // freely apart from their motor term
    h_bmotor_.assign(nabd_, 0);
    for(int b = 0; b < nabd_; b++)
    {
        h_bmotor_[b] = bt[bstart[b]] == 2;
        for(int v = bstart[b]; v < bstart[b + 1]; v++)
        {
            SG_CHECK((bt[v] == 2) == (h_bmotor_[b] != 0), SGIPC_EINVAL,
                     "an affine body is partly a motor");
            if(h_bmotor_[b])
                bt[v] = 0;
        }
        any_motor_ = any_motor_ || h_bmotor_[b];
    }
    for(int b = 0; b < nabd_; b++)
    {
        double ms = 0.0, c[3] = {0, 0, 0};
        for(int v = bstart[b]; v < bstart[b + 1]; v++)
        {
            ms += mass[v];
            for(int k = 0; k < 3; k++)
                c[k] += mass[v] * hx[3 * v + k];
            SG_CHECK((bt[v] != 0) == (bt[bstart[b]] != 0), SGIPC_EINVAL,
                     "an affine body is partly fixed");
        }
        SG_CHECK(ms > 0.0, SGIPC_EINVAL, "affine body without mass (no tets)");
        for(int k = 0; k < 3; k++)
            c[k] /= ms;
        bfix[b]   = bt[bstart[b]] != 0;
        double* M = bM.data() + 144 * (size_t)b;
        for(int v = bstart[b]; v < bstart[b + 1]; v++)
        {
            double* xb = xbar.data() + 3 * (size_t)v;
            for(int k = 0; k < 3; k++)
                xb[k] = hx[3 * v + k] - c[k];
            // J^T J = [[I, X], [X^T, blockdiag(xb xb^T)]] with X_k = e_k xb^T
            for(int k = 0; k < 3; k++)
            {
                M[12 * k + k] += mass[v];
                for(int l = 0; l < 3; l++)
                {
                    M[12 * k + 3 + 3 * k + l] += mass[v] * xb[l];
                    M[12 * (3 + 3 * k + l) + k] += mass[v] * xb[l];
                    for(int l2 = 0; l2 < 3; l2++)  // xb_l xb_l2 first: M exactly symmetric
                        M[12 * (3 + 3 * k + l) + 3 + 3 * k + l2] += mass[v] * (xb[l] * xb[l2]);
                }
            }
        }
        double* Mi = h_bMinv_.data() + 144 * (size_t)b;
        SG_CHECK(invert_dense(12, M, Mi), SGIPC_EINVAL, "singular affine-body mass (flat body)");
        for(int k = 0; k < 3; k++)
        {
            bq[12 * b + k]         = c[k];
            bq[12 * b + 3 + 4 * k] = 1.0;
        }
    }
    h_bstart_ = bstart;
    h_xbar_   = xbar;
    h_mass_   = mass;
    h_bfix_   = bfix;
    for(int b = 0; b < nabd_; b++)
        if(!bfix[b])
        {
            std::vector<double> gv(3 * (size_t)nv_, 0.0);
            for(int v = bstart[b]; v < bstart[b + 1]; v++)
                for(int k = 0; k < 3; k++)
                    gv[3 * v + k] = d.gravity[k];
            body_fit(b, gv.data(), bG.data() + 12 * b);
            body_fit(b, hv.data(), bqv.data() + 12 * b);
        }
    h_bG_ = bG;
    // ABD vertices follow their body: x = J q, v = J q', x~ = J (q + dt q' + dt^2 G)
    std::vector<double> hxt(3 * (size_t)nv_);
    for(int i = 0; i < nfem_; i++)
        for(int c = 0; c < 3; c++)
            hxt[3 * i + c] = hx[3 * i + c]
                             + (hv[3 * i + c] * d.dt + (bt[i] == 0 ? d.gravity[c] * (d.dt * d.dt) : 0.0));
    for(int b = 0; b < nabd_; b++)
    {
        double qt[12];
        for(int k = 0; k < 12; k++)
            qt[k] = bq[12 * b + k] + (bqv[12 * b + k] * d.dt + bG[12 * b + k] * (d.dt * d.dt));
        body_apply(b, bq.data() + 12 * b, hx.data());
        body_apply(b, bqv.data() + 12 * b, hv.data());
        body_apply(b, qt, hxt.data());
    }

    // ---- cloth constants: rest-frame Dm^-1 (2x2), area, lumped mass
    // area * thickness * density / 3 per vertex, interior-edge hinges with
    // their rest angle and length (femEnergy.cu:262-309, load_mesh.cpp:320-379;
    // the reference's initFEM leaves the cloth set-up out, main.cu:351-352)
    ntri_ = d.n_tris;
    SG_CHECK(ntri_ == 0 || d.tris, SGIPC_EINVAL, "cloth triangles missing");
    for(int e = 0; e < 3 * ntri_; e++)
        SG_CHECK(d.tris[e] < (uint32_t)nfem_, SGIPC_EINVAL,
                 "cloth vertex index out of range (or on an affine body)");
    std::vector<double>   tdm((size_t)4 * ntri_), tarea(ntri_);
    std::vector<uint32_t> hinges;
    std::vector<double>   th0, len0;
    for(int f = 0; f < ntri_; f++)
    {
        const uint32_t* t = d.tris + 3 * (size_t)f;
        double          u[3], v[3], n[3];
        for(int c = 0; c < 3; c++)
        {
            u[c] = xr[3 * t[1] + c] - xr[3 * t[0] + c];
            v[c] = xr[3 * t[2] + c] - xr[3 * t[0] + c];
        }
        n[0]      = u[1] * v[2] - u[2] * v[1];
        n[1]      = u[2] * v[0] - u[0] * v[2];
        n[2]      = u[0] * v[1] - u[1] * v[0];
        double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        SG_CHECK(nn > 0.0, SGIPC_EINVAL, "degenerate rest cloth triangle");
        tarea[f] = 0.5 * nn;
        for(int c = 0; c < 3; c++)
            n[c] /= nn;
        // minimal rotation taking n to +z: I + [r]x + [r]x^2 / (1 + cos), r = n x z
        double r[3] = {n[1], -n[0], 0.0}, cs = n[2];
        double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        if(cs + 1.0 == 0.0)
        {
            R[0] = -1.0;
            R[4] = -1.0;
        }
        else
        {
            double K[9] = {0, -r[2], r[1], r[2], 0, -r[0], -r[1], r[0], 0};
            for(int i = 0; i < 3; i++)
                for(int j = 0; j < 3; j++)
                {
                    double k2 = K[i * 3] * K[j] + K[i * 3 + 1] * K[3 + j] + K[i * 3 + 2] * K[6 + j];
                    R[i * 3 + j] += K[i * 3 + j] + k2 * (1.0 / (1.0 + cs));
                }
        }
        double pr[3][2];
        for(int k = 0; k < 3; k++)
            for(int i = 0; i < 2; i++)
                pr[k][i] = R[i * 3] * xr[3 * t[k]] + R[i * 3 + 1] * xr[3 * t[k] + 1]
                           + R[i * 3 + 2] * xr[3 * t[k] + 2];
        double D[4] = {pr[1][0] - pr[0][0], pr[2][0] - pr[0][0], pr[1][1] - pr[0][1],
                       pr[2][1] - pr[0][1]};
        double det  = D[0] * D[3] - D[1] * D[2];
        double inv[4] = {D[3] / det, -D[1] / det, -D[2] / det, D[0] / det};
        for(int k = 0; k < 4; k++)
            tdm[(size_t)k * ntri_ + f] = inv[k];
        double m = tarea[f] * d.cloth_thickness * d.cloth_density / 3.0;
        for(int k = 0; k < 3; k++)
            mass[t[k]] += m;
    }
    {
        // half-edges (a < b, opposite vertex, order) sorted: interior edges
        // are the keys seen exactly twice, opposite vertices in triangle order
        std::vector<std::array<uint32_t, 4>> he;
        he.reserve(3 * (size_t)ntri_);
        for(int f = 0; f < ntri_; f++)
            for(int k = 0; k < 3; k++)
            {
                uint32_t v0 = d.tris[3 * (size_t)f + k], v1 = d.tris[3 * (size_t)f + (k + 1) % 3];
                he.push_back({std::min(v0, v1), std::max(v0, v1), (uint32_t)(3 * f + k),
                              d.tris[3 * (size_t)f + (k + 2) % 3]});
            }
        std::sort(he.begin(), he.end());
        for(size_t i = 0; i < he.size();)
        {
            size_t j = i;
            while(j < he.size() && he[j][0] == he[i][0] && he[j][1] == he[i][1])
                j++;
            if(j - i == 2)
            {
                uint32_t q[4] = {he[i][0], he[i][1], he[i][3], he[i + 1][3]};
                hinges.insert(hinges.end(), q, q + 4);
                const double* P[4] = {xr + 3 * q[0], xr + 3 * q[1], xr + 3 * q[2], xr + 3 * q[3]};
                th0.push_back(hinge_theta(P[0], P[1], P[2], P[3]));
                double l2 = 0;
                for(int c = 0; c < 3; c++)
                    l2 += (P[1][c] - P[0][c]) * (P[1][c] - P[0][c]);
                len0.push_back(std::sqrt(l2));
            }
            i = j;
        }
    }
    nh_ = (int)th0.size();

    // ---- static BSR pattern: element adjacency (tets, cloth triangles,
    // hinges) + every diagonal
    std::vector<std::vector<int>> adj(nn_);
    for(int v = 0; v < nn_; v++)
        adj[v].push_back(v);
    auto link = [&](const uint32_t* ids, int m) {
        for(int a = 0; a < m; a++)
            for(int b = 0; b < m; b++)
                if(a != b)
                    adj[ids[a]].push_back((int)ids[b]);
    };
    for(int e = 0; e < nt_; e++)
        link(etets.data() + 4 * (size_t)e, 4);
    for(int b = 0; b < nabd_; b++)
        for(int m = 0; m < 4; m++)
            for(int n = 0; n < 4; n++)
                if(m != n)
                    adj[nfem_ + 4 * b + m].push_back(nfem_ + 4 * b + n);
    for(int f = 0; f < ntri_; f++)
        link(d.tris + 3 * (size_t)f, 3);
    for(int h = 0; h < nh_; h++)
        link(hinges.data() + 4 * (size_t)h, 4);
    std::vector<int> rp(nn_ + 1, 0), cl;
    for(int v = 0; v < nn_; v++)
    {
        auto& a = adj[v];
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
        rp[v + 1] = rp[v] + (int)a.size();
    }
    cl.reserve(rp[nn_]);
    for(int v = 0; v < nn_; v++)
        cl.insert(cl.end(), adj[v].begin(), adj[v].end());
    A_.set_pattern(nn_, rp, cl, s_);
    auto slot_of = [&](int r, int c) {
        auto beg = cl.begin() + rp[r], end = cl.begin() + rp[r + 1];
        return (int)(std::lower_bound(beg, end, c) - cl.begin());
    };

    // ---- element gather lists (k_fem_gather_hess / k_fem_gather_grad):
    // every upper block (i <= j) of the static pattern lists the element
    // blocks that add into it (boundary vertices carry no elastic Hessian);
    // every vertex lists the element gradient entries that add into it.
    // Records: tets at 0 (kFemStride each), triangles at off_tri
    // (kTriStride), hinges at off_hinge (kHingeStride)
    size_t off_tri = (size_t)kFemStride * nt_;
    size_t off_hinge = off_tri + (size_t)kTriStride * ntri_;
    size_t buf_len   = off_hinge + (size_t)kHingeStride * nh_;
    SG_CHECK(buf_len / 3 < (size_t)1 << 30, SGIPC_EINVAL, "too many elements for the gather codes");
    {
        std::vector<long long> key;  // upper slot << 32 | code
        key.reserve((size_t)10 * nt_ + 6 * (size_t)ntri_ + 10 * (size_t)nh_);
        std::vector<std::pair<int, int>> vkey;  // (vertex, gradient code)
        vkey.reserve((size_t)4 * nt_ + 3 * (size_t)ntri_ + 4 * (size_t)nh_);
        auto add_elem = [&](const uint32_t* ids, int m, size_t base, int npairs, const int* PA,
                            const int* PB) {
            for(int pi = 0; pi < npairs; pi++)
            {
                int va = (int)ids[PA[pi]], vb = (int)ids[PB[pi]];
                if(bt[va] != 0 || bt[vb] != 0)
                    continue;
                bool tr   = va > vb;
                int  slot = tr ? slot_of(vb, va) : slot_of(va, vb);
                long long code = (long long)(((base + 9 * (size_t)pi) / 3) << 1) | (tr ? 1 : 0);
                key.push_back(((long long)slot << 32) | code);
            }
            for(int a = 0; a < m; a++)
                vkey.push_back({(int)ids[a], (int)((base + 9 * (size_t)npairs + 3 * (size_t)a) / 3)});
        };
        for(int e = 0; e < nt_; e++)
            add_elem(etets.data() + 4 * (size_t)e, 4, (size_t)kFemStride * e, 10, kPairA, kPairB);
        for(int f = 0; f < ntri_; f++)
            add_elem(d.tris + 3 * (size_t)f, 3, off_tri + (size_t)kTriStride * f, 6, kTriPairA,
                     kTriPairB);
        for(int h = 0; h < nh_; h++)
            add_elem(hinges.data() + 4 * (size_t)h, 4, off_hinge + (size_t)kHingeStride * h, 10,
                     kPairA, kPairB);
        std::sort(key.begin(), key.end());
        std::vector<int> up, lo, ptr, src;
        src.reserve(key.size());
        for(size_t k = 0; k < key.size(); k++)
        {
            int slot = (int)(key[k] >> 32);
            if(k == 0 || slot != (int)(key[k - 1] >> 32))
            {
                ptr.push_back((int)src.size());
                up.push_back(slot);
            }
            src.push_back((int)(key[k] & 0xffffffffll));
        }
        ptr.push_back((int)src.size());
        // transpose slot of each upper block
        std::vector<int> rowof(rp[nn_]);
        for(int v = 0; v < nn_; v++)
            for(int k = rp[v]; k < rp[v + 1]; k++)
                rowof[k] = v;
        lo.resize(up.size());
        for(size_t u = 0; u < up.size(); u++)
            lo[u] = slot_of(cl[up[u]], rowof[up[u]]);
        std::stable_sort(vkey.begin(), vkey.end(),
                         [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
                             return a.first < b.first;
                         });
        std::vector<int> vptr(nv_ + 1, 0), vsrc(vkey.size());
        for(size_t k = 0; k < vkey.size(); k++)
        {
            vptr[vkey[k].first + 1]++;
            vsrc[k] = vkey[k].second;
        }
        for(int v = 0; v < nv_; v++)
            vptr[v + 1] += vptr[v];
        gh_up_.upload(up, s_);
        gh_lo_.upload(lo, s_);
        gh_ptr_.upload(ptr, s_);
        gh_src_.upload(src, s_);
        gv_ptr_.upload(vptr, s_);
        gv_src_.upload(vsrc, s_);
        n_up_ = (int)up.size();
        fem_buf_.resize(buf_len > 0 ? buf_len : 1);
    }
    {
        std::vector<int> bslots(16 * (size_t)nabd_);
        for(int b = 0; b < nabd_; b++)
            for(int m = 0; m < 4; m++)
                for(int n = 0; n < 4; n++)
                    bslots[16 * (size_t)b + 4 * m + n] = slot_of(nfem_ + 4 * b + m, nfem_ + 4 * b + n);
        std::vector<int> node_of(nv_);
        for(int v = 0; v < nv_; v++)
            node_of[v] = v < nfem_ ? v : nfem_ + 4 * abd_of[v];
        bslots_.upload(bslots, s_);
        node_of_.upload(node_of, s_);
        xbar_.upload(xbar, s_);
        bstart_.upload(bstart, s_);
        bfix_.upload(bfix, s_);
        bM_.upload(bM, s_);
        bV_.upload(bvol, s_);
        bq_.upload(bq, s_);
        bqs_.upload(bq, s_);
        gbody_.resize(12 * (size_t)std::max(nabd_, 1));
        gbody_.zero(s_);
        AbdDev& a = dev_.abd;
        a.nb = nabd_, a.nfem = nfem_;
        a.bstart = bstart_.data(), a.fixed = bfix_.data(), a.M = bM_.data(), a.vol = bV_.data();
        a.q = bq_.data(), a.qs = bqs_.data(), a.gbody = gbody_.data(), a.slots = bslots_.data();
        a.kappa = d.abd_kappa;
        bmotor_.upload(h_bmotor_, s_);
        bqm_.resize(12 * (size_t)nabd_);
        bqm_.zero(s_);
        a.motor          = any_motor_ ? bmotor_.data() : nullptr;
        a.qm             = bqm_.data();
        a.motor_strength = d.motor_strength;
        dev_.nfem = nfem_, dev_.node_of = node_of_.data(), dev_.xbar = xbar_.data();
    }
    tri_.upload(d.tris, 3 * (size_t)ntri_, s_);
    tri_dminv_.upload(tdm, s_);
    tri_area_.upload(tarea, s_);
    hinge_.upload(hinges, s_);
    hinge_th0_.upload(th0, s_);
    hinge_len0_.upload(len0, s_);
    {
        ClothDev& c = dev_.cloth;
        c.nf = ntri_, c.nh = nh_;
        c.tris = tri_.data(), c.dminv = tri_dminv_.data(), c.area = tri_area_.data();
        c.hinges = hinge_.data(), c.theta0 = hinge_th0_.data(), c.len0 = hinge_len0_.data();
        double nu  = d.cloth_nu, h = d.cloth_thickness;
        c.stretch  = d.cloth_E / (2.0 * (1.0 + nu));           // main.cu:303-307
        c.bend     = d.cloth_bend_E * h * h * h / (24.0 * (1.0 - nu * nu));
        c.shear    = 0.03 * c.stretch * d.cloth_strain_rate;
        c.rate     = d.cloth_strain_rate;
        c.off_tri = off_tri, c.off_hinge = off_hinge;
    }
    double msum = 0;
    for(double m : mass)
        msum += m;
    mean_mass_ = msum / nv_;

    // ---- scene extents (GIPC::init, main.cu:853-863)
    double up[3] = {-1e32, -1e32, -1e32}, lo[3] = {1e32, 1e32, 1e32};
    for(int f = 0; f < nf_; f++)
        for(int k = 0; k < 3; k++)
            for(int c = 0; c < 3; c++)
            {
                double p = d.x[3 * d.faces[3 * f + k] + c];
                lo[c]    = p < lo[c] ? p : lo[c];
                up[c]    = up[c] < p ? p : up[c];
            }
    bbox_diag2_ = 0;
    for(int c = 0; c < 3; c++)
        bbox_diag2_ += (up[c] - lo[c]) * (up[c] - lo[c]);
    double u2[3] = {-1e32, -1e32, -1e32}, l2[3] = {1e32, 1e32, 1e32};
    bool   any   = false;
    for(int v = 0; v < nv_; v++)
        if(bt[v] == 0)
        {
            any = true;
            for(int c = 0; c < 3; c++)
            {
                double p = d.x[3 * v + c];
                l2[c]    = p < l2[c] ? p : l2[c];
                u2[c]    = u2[c] < p ? p : u2[c];
            }
        }
    scene_diag_ = 0;
    if(any)
        for(int c = 0; c < 3; c++)
            scene_diag_ += (u2[c] - l2[c]) * (u2[c] - l2[c]);
    scene_diag_ = std::sqrt(scene_diag_);
    if(d.bbox_diag > 0.0)
        bbox_diag2_ = d.bbox_diag * d.bbox_diag;
    if(d.scene_diag > 0.0)
        scene_diag_ = d.scene_diag;

    // ---- device state
    size_t n3 = 3 * (size_t)nv_;
    x_.upload(hx, s_);
    xr_.upload(xr, n3, s_);
    xprev_.upload(hx, s_);
    v_.upload(hv, s_);
    xt_.upload(hxt, s_);
    mass_.upload(mass, s_);
    dx_.resize(n3);
    dx_.zero(s_);
    xsave_.resize(n3);
    gc_.resize(n3);
    gs_.resize(n3);
    rhs_.resize(3 * (size_t)nn_);
    if(nabd_)
    {
        dy_.resize(3 * (size_t)nn_);
        dy_.zero(s_);
    }
    scratch_.resize(16);
    dminv_.upload(dminv, s_);
    vol_.upload(vol, s_);
    mu_.upload(mu, s_);
    lam_.upload(lam, s_);
    btype_.upload(bt, s_);
    body_.upload(body, s_);
    tets_.upload(etets, s_);
    faces_.upload(d.faces, 3 * (size_t)nf_, s_);
    edges_.upload(d.edges, 2 * (size_t)ne_, s_);
    surf_.upload(d.surf_verts, ns_, s_);
    gp_.resize(ns_ + 1);
    gcount_.resize(1);
    pcount_.resize(5);
    ccount_.resize(1);
    pair_cap_ = std::max(2 * (ns_ + ne_), 100000);
    cand_cap_ = std::max(4 * (ns_ + ne_), 200000);
    pairs_.resize(pair_cap_);
    pairs_raw_.resize(pair_cap_);
    cand_.resize(cand_cap_);
    overflow_.resize(4);
    overflow_.zero(s_);
    coo_.reserve(nn_, s_);

    dev_.nv = nv_, dev_.nt = nt_, dev_.ns = ns_, dev_.nf = nf_, dev_.ne = ne_;
    dev_.dt = d.dt, dev_.dHat = dHat_, dev_.goff = d.ground_offset;
    dev_.gn   = V3{d.ground_normal[0], d.ground_normal[1], d.ground_normal[2]};
    dev_.grav = V3{d.gravity[0], d.gravity[1], d.gravity[2]};
    dev_.use_ground = d.use_ground;
    dev_.x = x_.data(), dev_.xt = xt_.data(), dev_.mass = mass_.data();
    dev_.gshape = gs_.data(), dev_.gcontact = gc_.data(), dev_.vals = A_.vals.data();
    dev_.dminv = dminv_.data(), dev_.vol = vol_.data(), dev_.mu = mu_.data(), dev_.lam = lam_.data();
    dev_.btype = btype_.data(), dev_.body = body_.data(), dev_.diag = A_.diag.data();
    dev_.rowptr = A_.rowptr.data(), dev_.cols = A_.cols.data();
    dev_.fem_buf = fem_buf_.data(), dev_.n_up = n_up_;
    dev_.gh_up = gh_up_.data(), dev_.gh_lo = gh_lo_.data(), dev_.gh_ptr = gh_ptr_.data();
    dev_.gh_src = gh_src_.data(), dev_.gv_ptr = gv_ptr_.data(), dev_.gv_src = gv_src_.data();
    dev_.tets = tets_.data(), dev_.surf = surf_.data(), dev_.faces = faces_.data();
    dev_.edges = edges_.data();
    dev_.gp = gp_.data(), dev_.gcount = gcount_.data();
    dev_.pairs = pairs_.data(), dev_.pairs_raw = pairs_raw_.data(), dev_.pcount = pcount_.data();
    dev_.pair_cap = pair_cap_;
    dev_.overflow = overflow_.data();

    // lagged friction sets (built at every step start)
    fr_code_.resize(pair_cap_);
    fr_lam_.resize(pair_cap_);
    fr_coord_.resize(2 * (size_t)pair_cap_);
    fr_basis_.resize(6 * (size_t)pair_cap_);
    fr_count_.resize(2);
    fr_count_.zero(s_);
    fg_ids_.resize(ns_ + 1);
    fg_lam_.resize(ns_ + 1);
    FricDev& f = dev_.fric;
    f.mu       = d.friction_mu;
    f.mu_gd    = d.ground_friction_mu;
    f.eps2     = 1e-4 * bbox_diag2_ * d.dt * d.dt;  // fDhat dt^2 (GIPC.cu:8544, :9128)
    f.xprev    = xprev_.data();
    f.cap      = pair_cap_;
    f.code = fr_code_.data(), f.lam = fr_lam_.data(), f.coord = fr_coord_.data();
    f.basis = fr_basis_.data(), f.count = fr_count_.data();
    f.gids = fg_ids_.data(), f.glam = fg_lam_.data(), f.gcount = fr_count_.data() + 1;

    // animated soft targets: free FEM vertices, each at most once (main.cu
    // rejects a vertex selected by two constraints)
    nsoft_ = d.n_soft > 0 ? d.n_soft : 0;
    if(nsoft_)
    {
        SG_CHECK(d.soft_index && d.soft_target, SGIPC_EINVAL, "soft target arrays missing");
        std::vector<char> seen(nv_, 0);
        for(int i = 0; i < nsoft_; i++)
        {
            uint32_t v = d.soft_index[i];
            SG_CHECK(v < (uint32_t)nfem_, SGIPC_EINVAL,
                     "soft target vertex out of range or on an affine body");
            SG_CHECK(bt[v] == 0, SGIPC_EINVAL, "soft target on a fixed vertex");
            SG_CHECK(!seen[v], SGIPC_EINVAL, "a vertex carries two soft targets");
            seen[v] = 1;
        }
        soft_idx_.upload(d.soft_index, nsoft_, s_);
        soft_tgt_.upload(d.soft_target, 3 * (size_t)nsoft_, s_);
    }
    anim_sub_       = d.motion_rate > 0.0 ? 1.0 / d.motion_rate : 1.0;
    dev_.nsoft      = nsoft_;
    dev_.soft_idx   = soft_idx_.data();
    dev_.soft_tgt   = soft_tgt_.data();
    dev_.soft_k     = d.soft_stiffness;
    dev_.soft_rate  = 0.0;  // animation_fullRate before the first step (GIPC.cuh:114)

    EnergyLayout L = energy_layout(dev_);
    epart_.resize(L.off[kEnergyTerms - 1] + L.n[kEnergyTerms - 1] + 1);
    eterms_.resize(kEnergyTerms + 1);

    if(d.precond == SGIPC_PRECOND_MAS)
    {
        MasConfig mc;
        mc.fp32 = d.mas_precision == SGIPC_MAS_FP32;
        mc.blocked_invert = mas_blocked_invert();
        // MAS parts per node: FEM vertices keep theirs, each body is one part
        std::vector<int> npart(nn_);
        int              maxp = -1;
        for(int v = 0; v < nfem_; v++)
        {
            npart[v] = d.part_id ? d.part_id[v] : v / 16;
            maxp     = std::max(maxp, npart[v]);
        }
        for(int b = 0; b < nabd_; b++)
            for(int k = 0; k < 4; k++)
                npart[nfem_ + 4 * b + k] = maxp + 1 + b;
        mas_.init(A_, npart.data(), mc, s_);
    }
    mas_.prof = &prof_;
    pcg_.prof = &prof_;
    // block-Jacobi mode: each affine body's 12x12 block as one block
    // (ABDPreconditioner on the body range after the global 3x3 diagonal)
    if(nabd_ > 0)
        pcg_.set_bodies(nfem_, nabd_, bslots_.data());
    bf_.init(nf_, 3, faces_.data());
    be_.init(ne_, 2, edges_.data());
    build_bvh();
    build_cp();
    SG_HIP(hipStreamSynchronize(s_));
}

// q = M^-1 sum m J^T f over body b's vertices (least-squares affine fit)
void Context::body_fit(int b, const double* f, double* q) const
{
    double rhs[12] = {0};
    for(int v = h_bstart_[b]; v < h_bstart_[b + 1]; v++)
    {
        const double* xb = h_xbar_.data() + 3 * (size_t)v;
        for(int k = 0; k < 3; k++)
        {
            double mf = h_mass_[v] * f[3 * (size_t)v + k];
            rhs[k] += mf;
            for(int l = 0; l < 3; l++)
                rhs[3 + 3 * k + l] += mf * xb[l];
        }
    }
    const double* Mi = h_bMinv_.data() + 144 * (size_t)b;
    for(int i = 0; i < 12; i++)
    {
        double a = 0.0;
        for(int k = 0; k < 12; k++)
            a += Mi[12 * i + k] * rhs[k];
        q[i] = a;
    }
}

// f = J q on body b's vertices
void Context::body_apply(int b, const double* q, double* f) const
{
    for(int v = h_bstart_[b]; v < h_bstart_[b + 1]; v++)
    {
        const double* xb = h_xbar_.data() + 3 * (size_t)v;
        for(int k = 0; k < 3; k++)
            f[3 * (size_t)v + k] = q[k] + q[3 + 3 * k] * xb[0] + q[4 + 3 * k] * xb[1]
                                   + q[5 + 3 * k] * xb[2];
    }
}

Context::~Context()
{
    if(own_s_)
        (void)hipStreamDestroy(own_s_);
}

void Context::set_stream(hipStream_t s)
{
    SG_HIP(hipStreamSynchronize(s_));
    s_ = s ? s : own_s_;
}

double Context::read_d(const double* p)
{
    double h;
    SG_HIP(hipMemcpyAsync(&h, p, sizeof(double), hipMemcpyDeviceToHost, s_));
    SG_HIP(hipStreamSynchronize(s_));
    return h;
}

unsigned Context::read_u(const unsigned* p)
{
    unsigned h;
    SG_HIP(hipMemcpyAsync(&h, p, sizeof(unsigned), hipMemcpyDeviceToHost, s_));
    SG_HIP(hipStreamSynchronize(s_));
    return h;
}

void Context::check_overflow()
{
    int h[4];
    SG_HIP(hipMemcpyAsync(h, overflow_.data(), sizeof(h), hipMemcpyDeviceToHost, s_));
    SG_HIP(hipStreamSynchronize(s_));
    SG_CHECK(h[0] == 0 && h[1] == 0 && h[2] == 0, SGIPC_ECAPACITY,
             "contact buffer capacity exceeded (pairs/COO/CCD candidates)");
}

void Context::set_positions(const double* x)
{
    if(nabd_)
    {
        // affine-body vertices stay affine: x = J q with q fitted to x
        std::vector<double> hx(x, x + 3 * (size_t)nv_), q(12 * (size_t)nabd_);
        for(int b = 0; b < nabd_; b++)
        {
            body_fit(b, hx.data(), q.data() + 12 * b);
            body_apply(b, q.data() + 12 * b, hx.data());
        }
        bq_.upload(q, s_);
        x_.upload(hx, s_);
        SG_HIP(hipStreamSynchronize(s_));
        build_bvh();
        build_cp();
        SG_HIP(hipStreamSynchronize(s_));
        return;
    }
    x_.upload(x, 3 * (size_t)nv_, s_);
    build_bvh();
    build_cp();
    SG_HIP(hipStreamSynchronize(s_));
}

void Context::set_state(const double* x, const double* v, const double* xt, double kappa)
{
    size_t              n3 = 3 * (size_t)nv_;
    std::vector<double> hx(x, x + n3), hv(v, v + n3), hxt(xt, xt + n3);
    if(nabd_)
    {
        std::vector<double> q(12 * (size_t)nabd_);
        for(int b = 0; b < nabd_; b++)
        {
            double qv[12], qt[12];
            body_fit(b, hx.data(), q.data() + 12 * b);
            body_fit(b, hv.data(), qv);
            body_fit(b, hxt.data(), qt);
            body_apply(b, q.data() + 12 * b, hx.data());
            body_apply(b, qv, hv.data());
            body_apply(b, qt, hxt.data());
        }
        bq_.upload(q, s_);
        bqs_.upload(q, s_);
        dy_.zero(s_);
    }
    x_.upload(hx, s_);
    xprev_.upload(hx, s_);
    v_.upload(hv, s_);
    xt_.upload(hxt, s_);
    kappa_   = kappa;
    in_step_ = false;
    newton_k_ = 0;
    dx_.zero(s_);
    build_bvh();
    build_cp();
    SG_HIP(hipStreamSynchronize(s_));
}

void Context::set_pcg(const sgipc_pcg_cfg& c)
{
    SG_CHECK(c.rel_tol >= 0.0 && c.abs_tol >= 0.0, SGIPC_EINVAL, "negative PCG tolerance");
    SG_CHECK(c.solver_precision == SGIPC_SOLVE_FP64 || c.solver_precision == SGIPC_SOLVE_MIXED,
             SGIPC_EINVAL, "unknown solver precision");
    desc_.pcg_rel_tol                 = c.rel_tol;
    desc_.pcg_abs_tol                 = c.abs_tol;
    desc_.pcg_max_iter                = c.max_iter;
    desc_.pcg_use_preconditioned_norm = c.use_preconditioned_norm;
    desc_.solver_precision            = c.solver_precision;
    plain_iters_                      = 0.0;  // counts of the former settings
    pcg_.reset_precision();
}

void Context::get(const DBuf<double>& b, double* out)
{
    b.download(out, 3 * (size_t)nv_, s_);
    SG_HIP(hipStreamSynchronize(s_));
}

void Context::set_xtilde(const double* xt)
{
    xt_.upload(xt, 3 * (size_t)nv_, s_);
    SG_HIP(hipStreamSynchronize(s_));
}

// ------------------------------------------------------------------ phases
void Context::build_bvh()
{
    bf_.build(x_.data(), nullptr, 0.0, s_);
    be_.build(x_.data(), nullptr, 0.0, s_);
}

void Context::build_cp()
{
    launch_query_proximity(dev_, bf_, be_, xr_.data(), s_);
    launch_ground_detect(dev_, s_);
}

void Context::assemble(bool hessian)
{
    for(int attempt = 0;; attempt++)
    {
        if(hessian)
        {
            A_.vals.zero(s_);
            coo_.begin(s_);
        }
        dev_.hkeys  = coo_.hkeys.data();
        dev_.hvals  = coo_.hvals.data();
        dev_.hlist  = coo_.hlist.data();
        dev_.hcount = coo_.hcount.data();
        dev_.hfail  = coo_.hfail.data();
        dev_.hmask  = (unsigned long long)(coo_.hcap - 1);
        launch_init_diag(dev_, s_);
        launch_soft(dev_, hessian, gs_.data(), s_);
        launch_abd_body(dev_, hessian, s_);
        launch_barrier(dev_, kappa_, hessian, gc_.data(), s_);
        launch_friction(dev_, hessian, gc_.data(), s_);
        launch_ground(dev_, kappa_, hessian, gc_.data(), s_);
        prof_.begin("fem", s_);
        launch_fem(dev_, hessian, s_);
        prof_.end("fem", s_);
        launch_rhs(dev_, rhs_.data(), s_);
        launch_abd_rhs(dev_, rhs_.data(), s_);
        if(!hessian)
            return;
        if(coo_.finish(s_))
            break;
        // the contact table ran full: it has been enlarged, assemble again
        SG_CHECK(attempt < 6, SGIPC_ECAPACITY, "contact block table cannot grow further");
    }
    coo_.build_rows(s_);
}

double Context::energy()
{
    launch_energy(dev_, kappa_, epart_.data(), eterms_.data(), s_);
    double t[kEnergyTerms + 1];
    SG_HIP(hipMemcpyAsync(t, eterms_.data(), sizeof(t), hipMemcpyDeviceToHost, s_));
    SG_HIP(hipStreamSynchronize(s_));
    return t[kEnergyTerms];
}

bool Context::is_intersected()
{
    int* hit = overflow_.data() + 3;
    SG_HIP(hipMemsetAsync(hit, 0, sizeof(int), s_));
    if(desc_.use_ground)
        launch_ground_intersect(dev_, hit, s_);
    int h = 0;
    SG_HIP(hipMemcpyAsync(&h, hit, sizeof(int), hipMemcpyDeviceToHost, s_));
    SG_HIP(hipStreamSynchronize(s_));
    if(h)
        return true;
    launch_edge_tri(dev_, be_, hit, s_);
    SG_HIP(hipMemcpyAsync(&h, hit, sizeof(int), hipMemcpyDeviceToHost, s_));
    SG_HIP(hipStreamSynchronize(s_));
    return h != 0;
}

double Context::ground_bound(double slack)
{
    if(ns_ < 1)
        return 1.0;
    SG_HIP(hipMemsetAsync(scratch_.data(), 0, sizeof(double), s_));
    launch_ground_ccd(dev_, dx_.data(), slack, scratch_.data(), s_);
    return 1.0 / read_d(scratch_.data());
}

double Context::self_bound(const int4* cand, const unsigned* count, int cap, double slack)
{
    unsigned n = read_u(count);
    if(n < 1)
        return 1.0;
    SG_HIP(hipMemsetAsync(scratch_.data() + 1, 0, sizeof(double), s_));
    launch_self_ccd(x_.data(), dx_.data(), cand, count, std::min<int>(cap, (int)n), slack,
                    scratch_.data() + 1, s_);
    return 1.0 / read_d(scratch_.data() + 1);
}

int Context::full_ccd(double alpha)
{
    bf_.build(x_.data(), dx_.data(), alpha, s_);
    be_.build(x_.data(), dx_.data(), alpha, s_);
    launch_query_ccd(dev_, bf_, be_, dx_.data(), alpha, cand_.data(), ccount_.data(), cand_cap_,
                     overflow_.data() + 2, s_);
    check_overflow();
    return (int)read_u(ccount_.data());
}

double Context::cfl_speed()
{
    SG_HIP(hipMemsetAsync(scratch_.data() + 2, 0, sizeof(double), s_));
    launch_cfl(dev_, dx_.data(), scratch_.data() + 2, s_);
    return read_d(scratch_.data() + 2);
}

// x = x_s - alpha dx (FEM), q = q_s - alpha dq and x = J q (affine bodies)
void Context::step_forward(double alpha)
{
    launch_step_forward(dev_, xsave_.data(), dx_.data(), alpha, s_);
    if(nabd_)
        launch_abd_step(dev_, dy_.data(), alpha, s_);
}

void Context::line_search(double& alpha, double cfl_alpha, int* nls, double* e0, double* e1)
{
    double E0 = energy();
    if(e0)
        *e0 = E0;
    SG_HIP(hipMemsetAsync(scratch_.data() + 3, 0, sizeof(double), s_));
    launch_dot(3 * nv_, gc_.data(), dx_.data(), scratch_.data() + 3, s_);
    double c1m = desc_.armijo_c1 * read_d(scratch_.data() + 3);
    SG_HIP(hipMemcpyAsync(xsave_.data(), x_.data(), 3 * (size_t)nv_ * sizeof(double),
                          hipMemcpyDeviceToDevice, s_));
    if(nabd_)
        SG_HIP(hipMemcpyAsync(bqs_.data(), bq_.data(), 12 * (size_t)nabd_ * sizeof(double),
                              hipMemcpyDeviceToDevice, s_));
    step_forward(alpha);
    build_bvh();
    // the halvings end at alpha = 0 (the intersection-free start) unless the
    // direction is not finite; bounded so that case reports an error
    for(int h = 0; is_intersected(); h++)
    {
        SG_CHECK(h < 1100, SGIPC_ENUMERIC, "line search found no intersection-free step");
        alpha /= 2.0;
        alpha = std::min(cfl_alpha, alpha);
        step_forward(alpha);
        build_bvh();
    }
    build_cp();
    double E  = energy();
    double LF = alpha;
    int    k  = 0;
    for(int i = 0; i < 30; i++)
    {
        if(E <= E0 + c1m * alpha)
            break;
        alpha *= desc_.armijo_beta;
        k++;
        step_forward(alpha);
        build_bvh();
        build_cp();
        E = energy();
    }
    if(alpha < LF)
    {
        bool recompute = false;
        for(int h = 0; is_intersected(); h++)
        {
            SG_CHECK(h < 1100, SGIPC_ENUMERIC, "line search found no intersection-free step");
            alpha *= desc_.armijo_beta;
            alpha = std::clamp(alpha, desc_.armijo_alpha_min, cfl_alpha);
            step_forward(alpha);
            build_bvh();
            recompute = true;
        }
        if(recompute)
            build_cp();
    }
    check_overflow();
    if(nls)
        *nls = k;
    if(e1)
        *e1 = E;
}

static double compute_H_b(double d, double dHat)
{
    double t = d - dHat;
    return (std::log(d / dHat) * -2.0 - t * 4.0 / d) + 1.0 / (d * d) * (t * t);
}

void Context::suggest_kappa(double& k)
{
    double H_b = compute_H_b(1.0e-16 * bbox_diag2_, dHat_);
    k = (mean_mass_ == 0.0) ? 1e11 / (4.0e-16 * bbox_diag2_ * H_b)
                            : 1e11 * mean_mass_ / (4.0e-16 * bbox_diag2_ * H_b);
}

void Context::upper_bound_kappa(double& k)
{
    double H_b  = compute_H_b(1.0e-16 * bbox_diag2_, dHat_);
    double kmax = 100 * 1e11 * mean_mass_ / (4.0e-16 * bbox_diag2_ * H_b);
    if(mean_mass_ == 0.0)
        kmax = 100 * 1e11 / (4.0e-16 * bbox_diag2_ * H_b);
    if(k > kmax)
        k = kmax;
}

void Context::init_kappa()
{
    if(read_u(pcount_.data()) == 0)
        return;
    // GE = kinetic + tet elastic gradient (gs; the reference leaves the cloth
    // gradient out, GIPC.cu:9809), gc = unit-kappa ground + barrier
    launch_init_diag(dev_, s_);
    launch_fem(dev_, false, s_, false);
    launch_soft(dev_, false, gs_.data(), s_);  // computeSoftConstraintGradient(_GE)
    launch_ground(dev_, 1.0, false, gc_.data(), s_);
    launch_barrier(dev_, 1.0, false, gc_.data(), s_);
    SG_HIP(hipMemsetAsync(scratch_.data() + 4, 0, 2 * sizeof(double), s_));
    launch_dot(3 * nv_, gc_.data(), gs_.data(), scratch_.data() + 4, s_);
    launch_dot(3 * nv_, gc_.data(), gc_.data(), scratch_.data() + 5, s_);
    double h[2];
    SG_HIP(hipMemcpyAsync(h, scratch_.data() + 4, sizeof(h), hipMemcpyDeviceToHost, s_));
    SG_HIP(hipStreamSynchronize(s_));
    double minKappa = -h[0] / h[1];
    if(minKappa > 0.0)
        kappa_ = minKappa;
    suggest_kappa(minKappa);
    if(kappa_ < minKappa)
        kappa_ = minKappa;
    upper_bound_kappa(kappa_);
}

// Step-size bounds of the direction dx_ (solve_subIP, GIPC.cu:10786-10822):
// ground bound (slackness 0.8), ACCD over the proximity pairs, full-CCD
// candidates at that step and, when there are any, the CFL bound
// sqrt(dHat) / max|dx| * 0.5, re-bounded by ACCD over the candidates when the
// CCD step exceeds twice the CFL step.  out = {ground, self, candidates, max
// speed (0 without candidates), alpha_cfl, alpha}.
void Context::step_bounds(double out[6])
{
    double alpha = 1.0, ag = 1.0;
    if(desc_.use_ground)
        ag = ground_bound(0.8);
    alpha     = std::min(alpha, ag);
    double as = self_bound(pairs_raw_.data(), pcount_.data(), pair_cap_, 0.8);
    alpha     = std::min(alpha, as);
    double temp_alpha = alpha, alpha_cfl = alpha, maxSpeed = 0.0;
    int    nccd       = full_ccd(temp_alpha);
    if(nccd > 0)
    {
        maxSpeed  = cfl_speed();
        alpha_cfl = std::sqrt(dHat_) / maxSpeed * 0.5;
        alpha     = std::min(alpha, alpha_cfl);
        if(temp_alpha > 2 * alpha_cfl)
        {
            alpha = std::min(temp_alpha, self_bound(cand_.data(), ccount_.data(), cand_cap_, 0.8));
            alpha = std::max(alpha, alpha_cfl);
        }
    }
    out[0] = ag;
    out[1] = as;
    out[2] = nccd;
    out[3] = maxSpeed;
    out[4] = alpha_cfl;
    out[5] = alpha;
}

void Context::ccd_bound(const double* dx, double out[6])
{
    dx_.upload(dx, 3 * (size_t)nv_, s_);
    step_bounds(out);
}

int Context::ccd_candidates(const double* dx, double alpha, int32_t* out, int cap)
{
    dx_.upload(dx, 3 * (size_t)nv_, s_);
    int n = full_ccd(alpha);
    int m = std::min(n, std::min(cap, cand_cap_));
    if(out && m > 0)
    {
        SG_HIP(hipMemcpyAsync(out, cand_.data(), sizeof(int4) * (size_t)m, hipMemcpyDeviceToHost,
                              s_));
        SG_HIP(hipStreamSynchronize(s_));
    }
    return n;
}

// ------------------------------------------------------------- iteration
void Context::newton_iteration(sgipc_iter_stats* st)
{
    sgipc_iter_stats loc{};
    hipEvent_t       ev[5];
    for(auto& e : ev)
        SG_HIP(hipEventCreate(&e));
    SG_HIP(hipEventRecord(ev[0], s_));
    assemble(true);
    check_overflow();
    SG_HIP(hipMemsetAsync(scratch_.data() + 6, 0, sizeof(double), s_));
    launch_max_abs(3 * nv_, dx_.data(), scratch_.data() + 6, s_);
    double xdelta = read_d(scratch_.data() + 6);
    SG_HIP(hipEventRecord(ev[1], s_));
    loc.xdelta   = xdelta;
    loc.n_pairs  = (int)read_u(pcount_.data());
    loc.n_ground = (int)read_u(gcount_.data());
    int k        = newton_k_;
    if(k == 1)
        xdelta0_ = xdelta;
    if(k > 0 && (xdelta < scene_diag_ * desc_.abs_xdelta_tol || xdelta < desc_.rel_xdelta_tol * xdelta0_))
    {
        loc.converged = 1;
        SG_HIP(hipStreamSynchronize(s_));
        loc.ms_assembly = loc.ms_total = now_ms(ev[0], ev[1]);
        for(auto& e : ev)
            (void)hipEventDestroy(e);
        if(st)
            *st = loc;
        return;
    }
    // the system passed to the solver, when an export is armed
    // (GlobalLinearSystem::solve_linear_system, global_linear_system.cu:388-405)
    if(mm_armed_)
    {
        std::vector<double> g, vl;
        std::vector<int>    rp, cl;
        gather_system(g, rp, cl, vl, false);
        std::string b_path, A_path = mm_system_paths(mm_dir_, mm_frame_, mm_ns_++, &b_path);
        mm_write_system(A_path, b_path, nn_, rp, cl, vl, g);
    }
    // ---- linear solve
    SysView S{&A_, &coo_};
    Precond pc = desc_.precond == SGIPC_PRECOND_MAS ? Precond::Mas : Precond::Diag;
    if(pc == Precond::Mas)
        mas_.build(S, s_);
    PcgConfig cfg;
    cfg.rel_tol      = desc_.pcg_rel_tol;
    cfg.abs_tol      = desc_.pcg_abs_tol;
    cfg.max_iter     = desc_.pcg_max_iter;
    cfg.use_rz       = desc_.pcg_use_preconditioned_norm != 0;
    cfg.mixed        = desc_.solver_precision == SGIPC_SOLVE_MIXED;
    double*  sol     = nabd_ ? dy_.data() : dx_.data();
    PcgStats ps      = solve_linear(S, pc, cfg, sol, &loc.warm_start, &loc.warm_fallback);
    SG_CHECK(ps.rel_residual == ps.rel_residual, SGIPC_ENUMERIC,
             "the linear solve produced a non-finite direction");
    if(nabd_)  // vertex displacements J dq of the affine-body vertices
        launch_abd_expand(dev_, dy_.data(), dx_.data(), s_);
    loc.pcg_iterations   = ps.iterations;
    loc.pcg_rel_residual = ps.rel_residual;
    loc.solver_fallback  = ps.outer_iterations > 1 ? 1 : 0;
    SG_HIP(hipEventRecord(ev[2], s_));
    // ---- step bounds
    double bnd[6];
    step_bounds(bnd);
    double alpha = bnd[5], alpha_cfl = bnd[4];
    loc.alpha_ccd = alpha;
    SG_HIP(hipEventRecord(ev[3], s_));
    // ---- line search
    line_search(alpha, alpha_cfl, &loc.line_search_steps, &loc.energy0, &loc.energy1);
    loc.alpha = alpha;
    if(kappa_ == 0.0)
        init_kappa();
    newton_k_++;
    SG_HIP(hipEventRecord(ev[4], s_));
    SG_HIP(hipStreamSynchronize(s_));
    loc.ms_assembly    = now_ms(ev[0], ev[1]);
    loc.ms_solve       = now_ms(ev[1], ev[2]);
    loc.ms_ccd         = now_ms(ev[2], ev[3]);
    loc.ms_line_search = now_ms(ev[3], ev[4]);
    loc.ms_total       = now_ms(ev[0], ev[4]);
    for(auto& e : ev)
        (void)hipEventDestroy(e);
    prof_.resolve();
    if(st)
        *st = loc;
}

// A sol = rhs.  With a warm-start mode on and a coarse MAS level, the solve
// continues from a coarse guess x0 (A y = rhs - A x0, sol = x0 + y) with the
// stop threshold of the zero-guess solve.  Plain MAS judges the guesses: the
// affine guess and, in mode 2, the network-corrected one are measured by
// their preconditioned residuals, the lower one is used when it beats the
// zero guess, and the zero-guess solve runs whenever the warm solve does not
// converge.  The warm solve is abandoned once it has run twice the PCG
// iterations that the recent zero-guess solves needed (their average with
// weight 1/2); its iterations are counted in the solve's total.
//
// Adaptive mode (the default): the guesses cost a Galerkin build, a coarse
// solve and one MAS apply and residual per guess, timed on the device.  From
// the chosen guess's preconditioned residual q_g against the zero guess's
// q_0, PCG's linear convergence predicts the fraction of iterations saved,
// f = log(q_0 / q_g) / log(q_0 / threshold), hence a saving of
// f / (1 - f) times the solve that ran (none when the zero guess won).  The
// net gain (saving minus cost) is averaged over the probes with weight 1/2;
// while that average is negative the warm start is paused for the next 2
// solves (8 when the probe saved less than a quarter of its cost), the pause
// doubling (up to 32) while it stays negative.  Occasional
// rejected guesses among paying ones therefore do not pause it.
PcgStats Context::solve_plain(const SysView& S, Precond pc, const PcgConfig& cfg, double* sol)
{
    PcgStats ps = pcg_.solve(S, pc, &mas_, rhs_.data(), sol, cfg, s_);
    if(ps.converged)
        plain_iters_ = plain_iters_ > 0.0 ? 0.5 * (plain_iters_ + ps.iterations) : ps.iterations;
    return ps;
}

PcgStats Context::solve_linear(const SysView& S, Precond pc, const PcgConfig& cfg, double* sol,
                               int* warm, int* warm_fallback)
{
    *warm          = 0;
    *warm_fallback = 0;
    if(learned_ && pc == Precond::Mas && warm_adaptive_ && warm_skip_ > 0)
    {
        warm_skip_--;
        return solve_plain(S, pc, cfg, sol);
    }
    if(learned_ && pc == Precond::Mas && coarse_.setup(mas_, s_))
    {
        hipEvent_t ev[3];
        for(auto& e : ev)
            SG_HIP(hipEventCreate(&e));
        SG_HIP(hipEventRecord(ev[0], s_));
        size_t n = 3 * (size_t)nn_;
        x0_.resize(n);
        b2_.resize(n);
        zt_.resize(n);
        // q[0] plain, q[1] affine, q[2] learned in the stop test's norm (r.z
        // or r.r); q[3 + k]: r.z of the same when the stop test uses r.r and
        // the adaptive mode judges (PCG's convergence is measured in r.z)
        double*    q      = scratch_.data() + 8;
        const bool zjudge = warm_adaptive_ && !cfg.use_rz;
        SG_HIP(hipMemsetAsync(q, 0, 6 * sizeof(double), s_));
        auto measure = [&](const double* r, int k) {
            if(cfg.use_rz || zjudge)
            {
                mas_.apply(r, zt_.data(), s_);
                launch_dot((int)n, r, zt_.data(), q + (cfg.use_rz ? k : 3 + k), s_);
            }
            if(!cfg.use_rz)
                launch_dot((int)n, r, r, q + k, s_);
        };
        coarse_.build(nn_, nfem_, btype_.data(), x_.data(), A_, &coo_, rhs_.data(), s_);
        coarse_.guess(nullptr, x0_.data(), s_);
        residual(S, rhs_.data(), x0_.data(), b2_.data(), s_);
        measure(rhs_.data(), 0);
        measure(b2_.data(), 1);
        bool net = learned_ == 2 && mlp_.ready();
        if(net)
        {
            x1_.resize(n);
            b3_.resize(n);
            coarse_.guess(&mlp_, x1_.data(), s_);
            residual(S, rhs_.data(), x1_.data(), b3_.data(), s_);
            measure(b3_.data(), 2);
        }
        SG_HIP(hipEventRecord(ev[1], s_));
        double h[6];
        SG_HIP(hipMemcpyAsync(h, q, sizeof(h), hipMemcpyDeviceToHost, s_));
        SG_HIP(hipStreamSynchronize(s_));
        int     pick = 1;  // the guess with the lowest preconditioned residual
        double  qb   = h[1];
        double* xg   = x0_.data();
        double* bg   = b2_.data();
        if(net && h[2] < h[1])
        {
            pick = 2;
            qb   = h[2];
            xg   = x1_.data();
            bg   = b3_.data();
        }
        // judge a probe: predicted saving against the measured cost
        auto judge = [&](double f) {
            SG_HIP(hipEventRecord(ev[2], s_));
            SG_HIP(hipEventSynchronize(ev[2]));
            double cost = now_ms(ev[0], ev[1]), solve = now_ms(ev[1], ev[2]);
            for(auto& e : ev)
                (void)hipEventDestroy(e);
            if(!warm_adaptive_)
                return;
            double saving = solve * f / (1.0 - f);
            double net    = saving - cost;
            warm_gain_    = warm_probes_++ ? 0.5 * (warm_gain_ + net) : net;
            if(warm_gain_ < 0.0)
            {
                const int first = saving < 0.25 * cost ? 8 : 2;
                warm_pause_     = warm_pause_ ? std::min(2 * warm_pause_, 32) : first;
                warm_skip_      = warm_pause_;
            }
            else
                warm_pause_ = 0;
        };
        int wasted = 0;
        if(h[0] > 0.0 && qb < h[0])
        {
            PcgConfig c = cfg;
            c.q_ref     = h[0];
            // abandon the guess at twice the iterations the recent solves
            // from zero needed
            if(plain_iters_ > 0.0)
            {
                int cap    = std::max(2 * (int)std::ceil(plain_iters_), 2 * cfg.check_every);
                c.max_iter = cfg.max_iter > 0 ? std::min(cfg.max_iter, cap) : cap;
            }
            PcgStats ps = pcg_.solve(S, pc, &mas_, bg, sol, c, s_);
            if(!ps.converged)
            {
                wasted         = ps.iterations;
                *warm_fallback = 1;
            }
            if(ps.converged)
            {
                launch_axpy_d((int)n, 1.0, xg, sol, s_);
                ps.rel_residual = pcg_.true_residual(S, rhs_.data(), sol, s_);
                *warm           = pick;
                double rel2 = cfg.rel_tol * cfg.rel_tol, abs2 = cfg.abs_tol * cfg.abs_tol;
                double thr  = std::max(rel2 * h[0], abs2);
                double z0 = zjudge ? h[3] : h[0], zg = zjudge ? h[3 + pick] : qb;
                double f  = (thr < h[0] && z0 > 0.0 && zg > 0.0)
                                ? std::log(z0 / zg) / std::log(h[0] / thr)
                                : 0.0;
                judge(std::min(std::max(f, 0.0), 0.95));
                return ps;
            }
        }
        PcgStats ps = solve_plain(S, pc, cfg, sol);
        ps.iterations += wasted;
        judge(0.0);
        return ps;
    }
    return solve_plain(S, pc, cfg, sol);
}

void Context::set_warm_adaptive(int on)
{
    warm_adaptive_ = on ? 1 : 0;
    warm_skip_     = 0;
    warm_pause_    = 0;
    warm_probes_   = 0;
}

void Context::set_learned(int mode)
{
    SG_CHECK(mode >= 0 && mode <= 2, SGIPC_EINVAL, "learned mode must be 0, 1 or 2");
    SG_CHECK(mode < 2 || mlp_.ready(), SGIPC_EINVAL, "no trained network (train or set it first)");
    learned_    = mode;
    warm_skip_   = 0;  // a new mode is probed afresh
    warm_pause_  = 0;
    warm_probes_ = 0;
}

int Context::learned_samples(float* feat, float* target, int cap)
{
    SG_CHECK(mas_.ready() && coarse_.setup(mas_, s_), SGIPC_EINVAL,
             "no coarse MAS level (run a solving Newton iteration with MAS first)");
    // positions of the solve: the line search started from xsave_
    coarse_.build(nn_, nfem_, btype_.data(), xsave_.data(), A_, &coo_, rhs_.data(), s_);
    return coarse_.samples(nabd_ ? dy_.data() : dx_.data(), feat, target, cap, s_);
}

void Context::begin_step()
{
    upper_bound_kappa(kappa_);
    if(kappa_ < 1e-16)
        suggest_kappa(kappa_);
    init_kappa();
    dev_.soft_rate = anim_sub_;  // animation_fullRate = animation_subRate (GIPC.cu:11018)
    if(any_motor_)
        set_motor_targets();
    // buildFrictionSets on the current contact sets (GIPC.cu:11016)
    launch_friction_build(dev_, kappa_, s_);
    dx_.zero(s_);
    if(nabd_)
        dy_.zero(s_);
    newton_k_ = 0;
    xdelta0_  = 0;
    in_step_  = true;
    step_log_.clear();
}

// IPC_Solver's animation loop (GIPC.cu:11038-11079): the step ends once
// animation_fullRate > 0.99; otherwise the rate advances by the sub-rate, the
// lagged friction sets are rebuilt and a fresh Newton loop (solve_subIP)
// starts from a zero direction
bool Context::next_newton_loop()
{
    if(dev_.soft_rate > 0.99)
        return true;
    dev_.soft_rate += anim_sub_;
    launch_friction_build(dev_, kappa_, s_);
    dx_.zero(s_);
    if(nabd_)
        dy_.zero(s_);
    newton_k_ = 0;
    xdelta0_  = 0;
    return false;
}

// Motor target of each motor body from its step prediction q~ (the affine
// fit of x~ over the body's vertices, exact since x~ = J q~): the reference's
// q_p = J^-1 J_delta q~ + q~ (J, J_delta the affine maps of the points 0, e_x,
// e_y, e_z and of their displacement under R_x(theta),
// setup_abd_system_gradient_and_hessian.cu:264-307) is A_p = A~ R_x(theta) on
// the rows a_i; rows a2 and a3 are then normalised.
void Context::set_motor_targets()
{
    std::vector<double> xt = xt_.to_host(s_);
    std::vector<double> qm(12 * (size_t)nabd_, 0.0);
    const double        th = desc_.motor_speed * desc_.dt;
    const double        c = std::cos(th), sn = std::sin(th);
    for(int b = 0; b < nabd_; b++)
    {
        if(!h_bmotor_[b])
            continue;
        double qt[12];
        body_fit(b, xt.data(), qt);
        double* qp = qm.data() + 12 * (size_t)b;
        for(int k = 0; k < 3; k++)
            qp[k] = 2.0 * qt[k];
        for(int i = 0; i < 3; i++)
        {
            const double* a     = qt + 3 + 3 * i;
            qp[3 + 3 * i]       = a[0];
            qp[3 + 3 * i + 1]   = c * a[1] + sn * a[2];
            qp[3 + 3 * i + 2]   = -sn * a[1] + c * a[2];
        }
        for(int i = 1; i < 3; i++)
        {
            double* a = qp + 3 + 3 * i;
            double  n = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
            for(int k = 0; k < 3; k++)
                a[k] /= n;
        }
    }
    bqm_.upload(qm, s_);
}

void Context::set_soft_targets(const double* t)
{
    SG_CHECK(nsoft_ > 0, SGIPC_ESTATE, "the scene has no soft targets");
    soft_tgt_.upload(t, 3 * (size_t)nsoft_, s_);
    SG_HIP(hipStreamSynchronize(s_));
}

void Context::end_step()
{
    mm_armed_ = false;
    launch_update_velocity(dev_, xprev_.data(), v_.data(), s_);
    in_step_ = false;
}

void Context::advance(int n, sgipc_advance_stats* st)
{
    hipEvent_t e0, e1;
    SG_HIP(hipEventCreate(&e0));
    SG_HIP(hipEventCreate(&e1));
    SG_HIP(hipEventRecord(e0, s_));
    sgipc_advance_stats loc{};
    int cap = desc_.newton_max_iter > 0 ? desc_.newton_max_iter : 10000;
    while(loc.newton_iterations < n)
    {
        if(!in_step_)
        {
            begin_step();
            loc.time_steps_begun++;
        }
        sgipc_iter_stats is{};
        newton_iteration(&is);
        step_log_.push_back(is);
        if(is.converged)
        {
            if(next_newton_loop())
                end_step();
            continue;
        }
        loc.newton_iterations++;
        loc.pcg_iterations += is.pcg_iterations;
        loc.solver_fallbacks += is.solver_fallback;
        loc.warm_starts += is.warm_start > 0 ? 1 : 0;
        loc.warm_fallbacks += is.warm_fallback;
        loc.max_pcg_rel_residual = std::max(loc.max_pcg_rel_residual, is.pcg_rel_residual);
        loc.ms_assembly += is.ms_assembly;
        loc.ms_solve += is.ms_solve;
        loc.ms_ccd += is.ms_ccd;
        loc.ms_line_search += is.ms_line_search;
        loc.line_search_steps += is.line_search_steps;
        loc.ccd_limited += is.alpha_ccd < 1.0 ? 1 : 0;
        if(newton_k_ >= cap)
        {
            loc.capped_steps++;
            if(next_newton_loop())
                end_step();
        }
    }
    SG_HIP(hipEventRecord(e1, s_));
    SG_HIP(hipStreamSynchronize(s_));
    loc.ms_total = now_ms(e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if(st)
        *st = loc;
}

void Context::profile_enable(int every)
{
    prof_.clear();
    prof_.enabled = every > 0;
    prof_.sample  = every > 0 ? every : 1;
}

void Context::profile_query(const char* name, int32_t* launches, double* total_ms)
{
    SG_HIP(hipStreamSynchronize(s_));
    prof_.resolve();
    auto e = prof_.query(name);
    if(launches)
        *launches = e.launches;
    if(total_ms)
        *total_ms = e.total_ms;
}

void Context::problem_sizes(int64_t* out)
{
    out[0] = nn_;
    out[1] = nt_;
    out[2] = A_.nnzb;
    out[3] = (int64_t)read_u(coo_.count.data());
    out[4] = mas_.ready() ? mas_.banks() : 0;
    out[5] = mas_.ready() ? mas_.levels() : 0;
    out[6] = (int64_t)read_u(pcount_.data());
    out[7] = (int64_t)read_u(gcount_.data());
}

void Context::time_step(sgipc_step_stats* st)
{
    hipEvent_t e0, e1;
    SG_HIP(hipEventCreate(&e0));
    SG_HIP(hipEventCreate(&e1));
    SG_HIP(hipEventRecord(e0, s_));
    begin_step();
    int  cap = desc_.newton_max_iter > 0 ? desc_.newton_max_iter : 10000;
    int  k = 0, pcg_total = 0;
    bool met = true;
    do  // one Newton loop per animation rate
    {
        bool loop_met = false;
        for(int i = 0; i < cap; i++)
        {
            sgipc_iter_stats is{};
            newton_iteration(&is);
            step_log_.push_back(is);
            if(is.converged)
            {
                loop_met = true;
                break;
            }
            pcg_total += is.pcg_iterations;
            k++;
        }
        met = met && loop_met;
    } while(!next_newton_loop());
    end_step();
    SG_HIP(hipEventRecord(e1, s_));
    SG_HIP(hipStreamSynchronize(s_));
    if(st)
    {
        st->newton_iterations = k;
        st->pcg_iterations    = pcg_total;
        st->ms_total          = now_ms(e0, e1);
        st->capped            = met ? 0 : 1;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

// ------------------------------------------------------------ ABI helpers
void Context::energy_terms(double* terms)
{
    launch_energy(dev_, kappa_, epart_.data(), eterms_.data(), s_);
    double t[kEnergyTerms + 1];
    SG_HIP(hipMemcpyAsync(t, eterms_.data(), sizeof(t), hipMemcpyDeviceToHost, s_));
    SG_HIP(hipStreamSynchronize(s_));
    terms[0] = t[kEnergyTerms];
    terms[1] = t[0];
    terms[2] = t[1];
    terms[3] = t[2];
    terms[4] = t[3];
    terms[5] = t[4] + t[5];
    terms[6] = t[6] + t[7];
    terms[7] = t[8];
    terms[8] = t[9];
}

void Context::lbvh_build(int which, uint64_t* keys, uint32_t* nodes, double* aabbs)
{
    Lbvh& b = which == 0 ? bf_ : be_;
    b.build(x_.data(), nullptr, 0.0, s_);
    int n = b.n;
    if(n > 0)
    {
        if(keys)
            b.keys.download(keys, n, s_);
        if(nodes)
            b.nodes.download(nodes, 4 * (size_t)(2 * n - 1), s_);
        if(aabbs)
            b.bvs.download(aabbs, 6 * (size_t)(2 * n - 1), s_);
    }
    SG_HIP(hipStreamSynchronize(s_));
}

void Context::contact_query(int32_t* pairs, int cap, int32_t* counts, uint32_t* ground, int gcap,
                            int32_t* n_ground)
{
    build_bvh();
    build_cp();
    check_overflow();
    unsigned c[5];
    SG_HIP(hipMemcpyAsync(c, pcount_.data(), sizeof(c), hipMemcpyDeviceToHost, s_));
    unsigned g = read_u(gcount_.data());
    if(counts)
        for(int i = 0; i < 5; i++)
            counts[i] = (int32_t)c[i];
    SG_CHECK(!pairs || (int)c[0] <= cap, SGIPC_ECAPACITY, "pair output buffer too small");
    SG_CHECK(!ground || (int)g <= gcap, SGIPC_ECAPACITY, "ground output buffer too small");
    if(pairs && c[0])
        SG_HIP(hipMemcpyAsync(pairs, pairs_.data(), c[0] * sizeof(int4), hipMemcpyDeviceToHost, s_));
    if(ground && g)
        SG_HIP(hipMemcpyAsync(ground, gp_.data(), g * sizeof(uint32_t), hipMemcpyDeviceToHost, s_));
    if(n_ground)
        *n_ground = (int32_t)g;
    SG_HIP(hipStreamSynchronize(s_));
}

// the assembled system on the host: right-hand side and full BSR (static
// pattern merged with the contact blocks), in the boundary DOF order.  With
// SGIPC_DOF_ABD_FIRST the body nodes come first (external node of internal
// node i: i - nfem for a body node, i + 4 nabd for a FEM vertex).
void Context::gather_system(std::vector<double>& grad, std::vector<int>& rp_out,
                            std::vector<int>& cl_out, std::vector<double>& vl_out,
                            bool reassemble)
{
    if(reassemble)
    {
        assemble(true);
        check_overflow();
    }
    grad.resize(3 * (size_t)nn_);
    rhs_.download(grad.data(), 3 * (size_t)nn_, s_);
    std::vector<double> v = A_.vals.to_host(s_);
    unsigned            nc = read_u(coo_.count.data());
    std::vector<int>    cr(nc), cc(nc);
    std::vector<double> cv(9 * (size_t)nc);
    if(nc)
    {
        coo_.r.download(cr.data(), nc, s_);
        coo_.c.download(cc.data(), nc, s_);
        coo_.v.download(cv.data(), 9 * (size_t)nc, s_);
    }
    SG_HIP(hipStreamSynchronize(s_));
    const bool abd_first = dof_order_ == SGIPC_DOF_ABD_FIRST && nabd_ > 0;
    auto ext = [&](int i) { return !abd_first ? i : (i >= nfem_ ? i - nfem_ : i + 4 * nabd_); };
    // merge static pattern and COO tail row by row (rows in boundary order)
    std::vector<std::vector<std::pair<int, const double*>>> extra(nn_);
    for(unsigned k = 0; k < nc; k++)
        extra[cr[k]].push_back({cc[k], cv.data() + 9 * (size_t)k});
    std::vector<std::vector<std::pair<int, std::array<double, 9>>>> rows(nn_);
    for(int i = 0; i < nn_; i++)
    {
        auto& row = rows[ext(i)];
        for(int k = A_.h_rowptr[i]; k < A_.h_rowptr[i + 1]; k++)
        {
            std::array<double, 9> b;
            for(int e = 0; e < 9; e++)
                b[e] = v[9 * (size_t)k + e];
            row.push_back({ext(A_.h_cols[k]), b});
        }
        for(auto& ex : extra[i])
        {
            bool found = false;
            for(auto& r : row)
                if(r.first == ext(ex.first))
                {
                    for(int e = 0; e < 9; e++)
                        r.second[e] += ex.second[e];
                    found = true;
                    break;
                }
            if(!found)
            {
                std::array<double, 9> b;
                for(int e = 0; e < 9; e++)
                    b[e] = ex.second[e];
                row.push_back({ext(ex.first), b});
            }
        }
    }
    rp_out.assign(nn_ + 1, 0);
    cl_out.clear();
    vl_out.clear();
    for(int i = 0; i < nn_; i++)
    {
        auto& row = rows[i];
        std::sort(row.begin(), row.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for(auto& r : row)
        {
            cl_out.push_back(r.first);
            vl_out.insert(vl_out.end(), r.second.begin(), r.second.end());
        }
        rp_out[i + 1] = (int)cl_out.size();
    }
    if(abd_first)
    {
        std::vector<double> g(grad.size());
        for(int i = 0; i < nn_; i++)
            for(int c = 0; c < 3; c++)
                g[3 * (size_t)ext(i) + c] = grad[3 * (size_t)i + c];
        grad.swap(g);
    }
}

void Context::assemble_out(double* grad, int32_t* nnzb, int32_t* rowptr, int32_t* cols,
                           double* vals, int cap)
{
    std::vector<double> g, vl;
    std::vector<int>    rp, cl;
    gather_system(g, rp, cl, vl);
    if(grad)
        std::copy(g.begin(), g.end(), grad);
    if(nnzb)
        *nnzb = (int32_t)cl.size();
    if(cap <= 0)
        return;
    SG_CHECK((int)cl.size() <= cap, SGIPC_ECAPACITY, "BSR output buffer too small");
    if(rowptr)
        std::copy(rp.begin(), rp.end(), rowptr);
    if(cols)
        std::copy(cl.begin(), cl.end(), cols);
    if(vals)
        std::copy(vl.begin(), vl.end(), vals);
}

void Context::mm_export_step(const std::string& dir, int frame)
{
    mm_dir_   = dir;
    mm_frame_ = frame;
    mm_ns_    = 0;
    mm_armed_ = true;
}

void Context::mm_export(const std::string& dir, int frame, int newton_step)
{
    std::vector<double> g, vl;
    std::vector<int>    rp, cl;
    gather_system(g, rp, cl, vl);
    std::string b_path;
    std::string A_path = mm_system_paths(dir, frame, newton_step, &b_path);
    mm_write_system(A_path, b_path, nn_, rp, cl, vl, g);
}

}  // namespace sgipc
