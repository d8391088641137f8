// sim_kernels.h -- plain-data view of the device state used by the element,
// contact and CCD kernels, plus their host launch wrappers.
#pragma once

#include "common.h"
#include "dmath.h"

namespace sgipc
{

// lagged friction sets (rebuilt at every time step start)
struct FricDev
{
    double    mu = 0.0, mu_gd = 0.0, eps2 = 0.0;
    const double* xprev = nullptr;  // positions at the step start
    int       cap    = 0;           // self pairs capacity
    int4*     code   = nullptr;     // lagged contact codes
    double*   lam    = nullptr;     // lagged normal force
    double*   coord  = nullptr;     // 2 per pair
    double*   basis  = nullptr;     // 6 per pair (3x2 row-major)
    unsigned* count  = nullptr;
    uint32_t* gids   = nullptr;     // ground contacts
    double*   glam   = nullptr;
    unsigned* gcount = nullptr;
};

// cloth triangles (membrane) and bending hinges
struct ClothDev
{
    int             nf = 0, nh = 0;
    const uint32_t* tris   = nullptr;  // 3 per triangle
    const double*   dminv  = nullptr;  // 2x2 rest Dm^-1, SoA: dminv[k * nf + f]
    const double*   area   = nullptr;
    const uint32_t* hinges = nullptr;  // 4 per hinge {a, b, opp0, opp1}
    const double*   theta0 = nullptr;  // rest dihedral angle
    const double*   len0   = nullptr;  // rest hinge-edge length
    double          stretch = 0, shear = 0, bend = 0, rate = 0;
    size_t          off_tri = 0, off_hinge = 0;  // record offsets in the element buffer
};

// affine bodies: vertices [nfem, nv) move with their body, x = p + A xbar;
// the system nodes are [FEM vertices | 4 per body (p, a1, a2, a3)]
struct AbdDev
{
    int           nb   = 0;        // bodies
    int           nfem = 0;        // first ABD vertex = first body node
    const int*    bstart = nullptr;  // nb + 1 vertex ranges
    const int*    fixed  = nullptr;  // per body
    const double* M      = nullptr;  // 12x12 dyadic mass per body (row-major)
    const double* vol    = nullptr;  // per body
    double*       q      = nullptr;  // 12 per body
    double*       qs     = nullptr;  // q at the line-search start
    double*       gbody  = nullptr;  // 12 per body: shape gradient
    const int*    slots  = nullptr;  // 16 per body: BSR slot of node block (m, n)
    double        kappa  = 0.0;
    // motor bodies: flag and target q_p per body, the motor's mass multiple
    const int*    motor  = nullptr;
    const double* qm     = nullptr;
    double        motor_strength = 0.0;
};

struct SimDev
{
    int    nv = 0, nt = 0, ns = 0, nf = 0, ne = 0;
    double dt = 0, dHat = 0, goff = 0;
    V3     gn{0, 1, 0}, grav{0, 0, 0};
    int    use_ground = 0;
    // state
    double *x = nullptr, *xt = nullptr, *mass = nullptr, *gshape = nullptr, *gcontact = nullptr;
    double* vals = nullptr;
    const double *dminv = nullptr, *vol = nullptr, *mu = nullptr, *lam = nullptr;
    const int *btype = nullptr, *body = nullptr, *diag = nullptr;
    // element assembly without atomics: k_fem writes each tet's 10 upper
    // blocks and 12 gradient entries (kFemStride doubles) to fem_buf, the
    // cloth kernels append their triangle and hinge records after the tets;
    // gathers add them per upper BSR block (gh_*) and per vertex (gv_*).
    // Gather codes address the buffer in units of 3 doubles: a Hessian code
    // is (offset / 3) << 1 | transposed, a gradient code is offset / 3
    double*    fem_buf = nullptr;
    int        n_up    = 0;
    const int *gh_up = nullptr, *gh_lo = nullptr, *gh_ptr = nullptr, *gh_src = nullptr;
    const int *gv_ptr = nullptr, *gv_src = nullptr;
    const int *rowptr = nullptr, *cols = nullptr;
    const uint32_t *tets = nullptr, *surf = nullptr, *faces = nullptr, *edges = nullptr;
    // ground contacts
    uint32_t* gp     = nullptr;
    unsigned* gcount = nullptr;
    // self contacts (int4 codes) and their raw candidates
    int4*     pairs    = nullptr;
    int4*     pairs_raw = nullptr;
    unsigned* pcount   = nullptr;
    int       pair_cap = 0;
    // COO tail
    // contact blocks outside the static pattern: open-addressing table keyed
    // by (row, col), merged as they are inserted (owned by Coo)
    unsigned long long* hkeys  = nullptr;
    double*             hvals  = nullptr;
    int*                hlist  = nullptr;
    unsigned*           hcount = nullptr;
    int*                hfail  = nullptr;
    unsigned long long  hmask  = 0;
    int*      overflow  = nullptr;  // [0] pairs, [1] coo, [2] ccd
    FricDev   fric;
    ClothDev  cloth;
    // node map of every vertex: node_of[v] = v (FEM) or the body's first
    // node; xbar = body-frame coordinates of ABD vertices (zero for FEM)
    int           nfem    = 0;
    const int*    node_of = nullptr;
    const double* xbar    = nullptr;
    AbdDev        abd;
    // animated soft targets (free FEM vertices, each at most once), the
    // motion stiffness and the current animation rate
    int             nsoft    = 0;
    const uint32_t* soft_idx = nullptr;
    const double*   soft_tgt = nullptr;
    double          soft_k = 0.0, soft_rate = 0.0;
};

constexpr int kFemStride = 102;  // 10 blocks x 9 + 12 gradient entries per tet
// (a, b), a <= b, of the 10 upper vertex pairs of a tet
constexpr int kPairA[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3};
constexpr int kPairB[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};

// partial-sum segments of the energy terms: kinetic, elastic, barrier,
// ground, friction, ground friction, cloth membrane, cloth bending, affine
// shape, soft targets
constexpr int kEnergyTerms = 10;
struct EnergyLayout
{
    int n[kEnergyTerms];
    int off[kEnergyTerms];
};

// fem.hip
void launch_init_diag(const SimDev& d, hipStream_t s);
// soft-target gradient into gshape and, with hessian, r^2 k I on the diagonal blocks
void launch_soft(const SimDev& d, bool hessian, double* gshape, hipStream_t s);
// cloth = false leaves the cloth records out of the gathered gradient
void launch_fem(const SimDev& d, bool hessian, hipStream_t s, bool cloth = true);
void launch_ground_detect(const SimDev& d, hipStream_t s);
void launch_ground(const SimDev& d, double kappa, bool hessian, double* gcontact, hipStream_t s);
void launch_barrier(const SimDev& d, double kappa, bool hessian, double* gcontact, hipStream_t s);
// friction.hip
void launch_friction_build(const SimDev& d, double kappa, hipStream_t s);
void launch_friction(const SimDev& d, bool hessian, double* gcontact, hipStream_t s);
void launch_energy_friction(const SimDev& d, const EnergyLayout& L, double* partial,
                            hipStream_t s);
// abd.hip
void launch_abd_body(const SimDev& d, bool hessian, hipStream_t s);
void launch_abd_rhs(const SimDev& d, double* rhs, hipStream_t s);
void launch_abd_expand(const SimDev& d, const double* dy, double* dx, hipStream_t s);
void launch_abd_step(const SimDev& d, const double* dy, double alpha, hipStream_t s);
void launch_energy_abd(const SimDev& d, const EnergyLayout& L, double* partial, hipStream_t s);
// cloth.hip
void launch_cloth(const SimDev& d, bool hessian, hipStream_t s);
void launch_energy_cloth(const SimDev& d, const EnergyLayout& L, double* partial, hipStream_t s);
void launch_rhs(const SimDev& d, double* rhs, hipStream_t s);
EnergyLayout energy_layout(const SimDev& d);
void launch_energy(const SimDev& d, double kappa, double* partial, double* terms, hipStream_t s);
void launch_dot(int n, const double* a, const double* b, double* out, hipStream_t s);
void launch_max_abs(int n, const double* a, double* out, hipStream_t s);
void launch_step_forward(const SimDev& d, const double* xs, const double* dx, double alpha,
                         hipStream_t s);
void launch_update_velocity(const SimDev& d, double* xprev, double* v, hipStream_t s);

// lbvh.hip: stable radix sort of 64-bit keys on bits [lo, hi) (8-bit digits);
// hist holds radix_sort_hist_size(n) counters
void   radix_sort_keys(uint64_t* keys, uint64_t* alt, unsigned* hist, int n, int lo, int hi,
                       hipStream_t s);
size_t radix_sort_hist_size(int n);
// ... of (key, 32-bit value) pairs: the values move with their keys (stable)
void radix_sort_pairs(uint64_t* keys, uint64_t* kalt, uint32_t* vals, uint32_t* valt,
                      unsigned* hist, int n, int lo, int hi, hipStream_t s);
// the digit-rounded bit count that covers keys up to max_key
int radix_key_bits(uint64_t max_key);

// mas_galerkin.hip: exclusive scan of n unsigned values into out (out may be
// in); tsum holds scan_tmp_size(n) entries, the total ends in
// tsum[scan_tmp_size(n) - 1]
void   scan_exclusive_u32(int n, const unsigned* in, unsigned* out, unsigned* tsum, hipStream_t s);
size_t scan_tmp_size(int n);

// spatial_sort.hip: sorted spatial keys and orders of vertices (kind 0) or tets (1)
void spatial_sort_keys(int kind, const double* x, int nv, const uint32_t* tets, int nt,
                       uint64_t* keys_out, uint32_t* order_out, hipStream_t s);

// lbvh.hip: linear BVH over faces (arity 3) or edges (arity 2)
struct Lbvh
{
    int               n = 0, arity = 0;
    const uint32_t*   prims = nullptr;
    DBuf<double>      leaf;    // 6 n
    DBuf<double>      bvs;     // 6 (2n - 1) {upper, lower}
    DBuf<uint32_t>    nodes;   // 4 (2n - 1) {parent, left, right, element}
    DBuf<uint64_t>    keys, keys_alt;
    DBuf<unsigned>    flags;
    DBuf<int>         rng;                // 2 (n - 1): sorted leaf range of each internal node
    DBuf<double>      agg_blk, agg_sup;   // unions of 64 leaves / of 64 such blocks
    DBuf<unsigned>    hist;
    DBuf<double>      scene, partial;
    void init(int n, int arity, const uint32_t* prims);
    // swept boxes when dx != nullptr: box(x) U box(x - alpha dx)
    void build(const double* x, const double* dx, double alpha, hipStream_t s);
};

// contact.hip

void launch_query_proximity(const SimDev& d, const Lbvh& bf, const Lbvh& be, const double* xrest,
                            hipStream_t s);
// swept candidates into cand (count in *ccount, capacity cap)
void launch_query_ccd(const SimDev& d, const Lbvh& bf, const Lbvh& be, const double* dx,
                      double alpha, int4* cand, unsigned* ccount, int cap, int* overflow,
                      hipStream_t s);
void launch_edge_tri(const SimDev& d, const Lbvh& be, int* hit, hipStream_t s);
void launch_ground_intersect(const SimDev& d, int* hit, hipStream_t s);

// ccd.hip: reductions write max(1/toc) style values into out (positive max)
void launch_self_ccd(const double* x, const double* dx, const int4* cand, const unsigned* count,
                     int cap, double slackness, double* out_max, hipStream_t s);
void launch_ground_ccd(const SimDev& d, const double* dx, double slackness, double* out_max,
                       hipStream_t s);
void launch_cfl(const SimDev& d, const double* dx, double* out_max, hipStream_t s);

}  // namespace sgipc
