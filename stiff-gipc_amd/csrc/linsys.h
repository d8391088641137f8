// linsys.h -- device linear system: full-storage BSR (CSR of 3x3 blocks) plus a
// dynamic COO tail for contact blocks outside the static pattern, the MAS
// preconditioner and PCG.
//
// Counterparts in the reference:
//   GIPCTripletMatrix + Converter (linear_system/linear_system/global_matrix.h,
//     linear_system/utils/converter.cu) -> static pattern built once from the
//     mesh, contact blocks appended to a COO tail (no per-iteration sort);
//   Spmv::warp_reduce_sym_spmv (linear_system/utils/spmv.cu:8-133) -> row-parallel
//     full-storage SpMV over the static pattern plus a row-sorted copy of the
//     contact tail, no atomics on y;
//   PCGSolver (linear_system/solver/pcg_solver.cu:155-304) -> Pcg below, with
//     device-resident scalars and one host sync per convergence test;
//   MASPreconditioner (MASPreconditioner.cu) -> Mas below.
#pragma once

#include <cstdlib>
#include <cstring>

#include "common.h"

namespace sgipc
{

// --------------------------------------------------------------- matrices
struct Bsr
{
    int          nb   = 0;  // block rows (= vertices)
    int          nnzb = 0;  // stored blocks, both triangles
    DBuf<int>    rowptr;    // nb + 1
    DBuf<int>    cols;      // nnzb, sorted within each row
    DBuf<int>    rows;      // nnzb, block row of each entry
    DBuf<int>    diag;      // nb, index of block (i, i)
    DBuf<double> vals;      // 9 * nnzb, row-major blocks
    DBuf<float>  vals32;    // FP32 copy for the mixed-precision inner solve
    std::vector<int> h_rowptr, h_cols;  // host copy of the pattern

    // build the pattern from host CSR (cols sorted, every row has its diagonal)
    void set_pattern(int nb, const std::vector<int>& rowptr, const std::vector<int>& cols,
                     hipStream_t s);
};

struct Coo
{
    // distinct contact blocks (row, col, 3x3), written by finish() from the
    // table the assembly inserts into
    int            cap = 0;  // capacity of r, c, v, rc, rv (entries)
    // host-side bound on the live entries (set by finish()): launches over
    // the entries are sized by it instead of cap
    unsigned       live = 0;
    DBuf<int>      r, c;     // cap
    DBuf<double>   v;        // 9 * cap
    DBuf<unsigned> count;    // 1 (device)
    // row-sorted copy (CSR over the block rows) made by build_rows()
    int            nb = 0;
    DBuf<int>      rp;       // nb + 1
    DBuf<int>      fill;     // nb + 1 (histogram / fill cursor)
    DBuf<int>      tiles;    // scan tile offsets
    DBuf<int>      rc;       // cap
    DBuf<double>   rv;       // 9 * cap
    DBuf<float>    rv32;     // FP32 copy (mixed precision)
    bool           rows_valid = false;
    // open-addressing table the contact kernels insert into (Coo::table_*):
    // keys (row << 32 | col, empty = ~0), 9 values per slot (zero between
    // assemblies), the list of claimed slots, its length and a full flag
    DBuf<unsigned long long> hkeys;
    DBuf<double>             hvals;
    DBuf<int>                hlist;
    DBuf<unsigned>           hcount;
    DBuf<int>                hfail;
    size_t                   hcap = 0;  // slots, a power of two
    void reserve(int nrows, hipStream_t s);
    // before an assembly: empty list, no entries
    void begin(hipStream_t s);
    // after an assembly: the claimed slots become the entries (count) and
    // are cleared; false when the table ran full (it has been enlarged and
    // the assembly must be redone)
    bool finish(hipStream_t s);
    // counting sort of the live entries by row (after every assembly)
    void build_rows(hipStream_t s);

  private:
    void grow_table(size_t slots, hipStream_t s);
    void ensure_entries(int n);
};

// view of A = static BSR + COO tail
struct SysView
{
    const Bsr* A   = nullptr;
    const Coo* coo = nullptr;  // may be null; its row-sorted copy must be current
};

// ------------------------------------------------------------ reductions
// (slot sets: common.h)
// host-side sum of a slot set in the device's butterfly order
double slot_set_sum_host(const double* h);
// after a launch of nblocks workgroups that called slot_set_add on set: add
// the sum of its partials to slot 0 (skipped, like the launch, while
// *done != 0)
void set_finalize(double* set, int nblocks, const int* done, hipStream_t s);
double read_slot_set(const double* d, hipStream_t s);

// y = A x (FP64); optional dot(x, y) accumulated into the slot set `dot`
void spmv(const SysView& S, const double* x, double* y, double* dot, hipStream_t s);
// r = b - A x
void residual(const SysView& S, const double* b, const double* x, double* r, hipStream_t s);
// y += a x (FP64)
void launch_axpy_d(int n, double a, const double* x, double* y, hipStream_t s);
// dot(a, b) accumulated into the slot set `out` (must be zeroed)
void dot_slots(int n, const double* a, const double* b, double* out, hipStream_t s);
// refresh the FP32 copies of the matrix (static pattern and contact rows)
void refresh_fp32(const SysView& S, hipStream_t s);

// ------------------------------------------------------------------ MAS
struct MasConfig
{
    bool fp32 = true;  // FP32 inverses and multilevel R/Z (reference default)
    // bank inverses: false = scalar Gauss-Jordan in the reference's order
    // (mas_invert.hip, bit-identical to a sequential restatement); true =
    // eliminated by 16x16 tiles with FP64 MFMA tile products (the same inverse
    // in exact arithmetic, rounding accumulated in another order)
    bool blocked_invert = false;
};

// SGIPC_MAS_INVERT=blocked selects the tiled MFMA bank inverse
inline bool mas_blocked_invert()
{
    const char* e = getenv("SGIPC_MAS_INVERT");
    return e && strcmp(e, "blocked") == 0;
}

void launch_mas_invert_gj(const double* B, float* inv, int nbanks, hipStream_t s);
void launch_mas_invert_gj(const double* B, double* inv, int nbanks, hipStream_t s);

// the bank matrices of MAS levels >= 1 from the Galerkin operators A_l,
// summed in a fixed order by sorting (mas_galerkin.hip)
class MasGalerkin
{
  public:
    void run(const Bsr& A, const Coo* coo, int L, const std::vector<int>& lvl, const int* chain,
             const int* going_next, double* B, hipStream_t s);

  private:
    void ensure(int n);
    int  reduce(int n, uint64_t kdead, int nsrc0, const double* src0, const double* src1,
                int offl, long long nl, double* B, int out, hipStream_t s);
    int            cap_ = 0;
    DBuf<uint64_t> keys_, kalt_;
    DBuf<uint32_t> vals_, valt_;
    DBuf<unsigned> flag_, pos_, hist_, tsum_;
    DBuf<int>      rstart_;  // start of each run of equal keys
    DBuf<int>      er_[2], ec_[2];
    DBuf<double>   eb_[2];
};

class Mas
{
  public:
    // static set-up: partition (non-decreasing part ids, <= 16 per part) and
    // the static adjacency of the pattern of A
    void init(const Bsr& A, const int* part_id_host, const MasConfig& cfg, hipStream_t s);
    // per Newton iteration: hierarchy (with COO connectivity), assembly and
    // block inverses (ReorderRealtime + PrepareHessian_bcoo + __inverse6_P96x96)
    void build(const SysView& S, hipStream_t s);
    // z = M^-1 r (standalone)
    void apply(const double* r, double* z, hipStream_t s);

    // fused PCG pieces (see Pcg): level-0 banks with the optional update
    // x += a p, r -= a Ap (a = rz[k-1] / pAp[k], both given as slot sets);
    // coarse levels; collect with rz accumulated into the slot set rz_out and
    // the two slot sets zero0/zero1 (may be null) cleared for later use.
    // Every kernel returns at once when done (may be null) is non-zero.
    // defer_rz: rz_out keeps collect's collect_blocks() workgroup partials
    // without the finalize launch (its reader sums them)
    void apply_fused(double* x, double* r, const double* p, const double* Ap,
                     const double* rz_prev, const double* pAp, double* z, double* rz_out,
                     double* zero0, double* zero1, const int* done, hipStream_t s,
                     bool defer_rz = false);
    // the same on FP32 PCG vectors (inner solve of the mixed-precision mode)
    void apply_fused(float* x, float* r, const float* p, const float* Ap,
                     const double* rz_prev, const double* pAp, float* z, double* rz_out,
                     double* zero0, double* zero1, const int* done, hipStream_t s,
                     bool defer_rz = false);
    // workgroups of the collect launch (partials of its rz set)
    int collect_blocks() const { return (int)grid_for(nv_, 256); }

    int levels() const { return L_; }
    int total_clusters() const { return total_clusters_; }
    const std::vector<int>& level_sizes() const { return h_lvl_; }  // (count, off) pairs
    std::vector<int> chain_host(hipStream_t s) const;
    bool ready() const { return built_; }
    const int*    chain_dev() const { return chain_.data(); }  // nv x (L - 1) global ids
    const double* bank_mats() const { return B_.data(); }      // nbanks x 48 x 48 Galerkin
    int  banks() const { return total_clusters_ / 16; }
    // assembled 48 x 48 matrix and inverse of one bank (any level), FP64 on
    // the host (an FP32 inverse is widened)
    void bank_host(int bank, double* mat, double* inv, hipStream_t s) const;
    Profiler* prof = nullptr;

  private:
    template <typename V>
    void apply_fused_t(V* x, V* r, const V* p, const V* Ap, const double* rz_prev,
                       const double* pap, V* z, double* rz_out, double* zero0, double* zero1,
                       const int* done, hipStream_t s, bool defer_rz);
    void cluster_level(int level, hipStream_t s);
    void assemble(const SysView& S, hipStream_t s);
    void invert(hipStream_t s);

    MasConfig cfg_;
    int       nv_ = 0, nparts_ = 0, total_map_ = 0, L_ = 0, total_clusters_ = 0;
    bool      built_ = false;
    DBuf<int>      slot_of_, vert_of_slot_;
    DBuf<unsigned> mask0_static_;
    DBuf<int>      xa_, xb_;   // static cross-bank edges (unordered)
    int            nx_ = 0;
    DBuf<unsigned> mask_;      // per-level scratch masks
    DBuf<unsigned> comp_;      // component masks
    DBuf<int>      bank_cnt_, bank_prefix_, newid_;
    DBuf<int>      node_;      // L x nv local node ids (levels >= 1)
    DBuf<int>      chain_;     // nv x (L-1) global ids
    DBuf<int>      lvl_;       // 2 * (L + 1): count, offset
    std::vector<int> h_lvl_;
    DBuf<double>   B_;         // nbanks x 48 x 48 assembled (FP64)
    DBuf<float>    inv32_;
    DBuf<double>   inv64_;
    DBuf<float>    R32_, Z32_;
    DBuf<double>   R64_, Z64_;
    DBuf<int>      going_next_; // global id -> global id one level up
    DBuf<int>      bank_first_; // global bank -> first next-level global id
    DBuf<int>      bank_ncl_;   // global bank -> number of next-level nodes
    DBuf<int>      child_bank_; // global id (levels >= 1) -> bank holding its members
    size_t         nbanks_alloc_ = 0;
    MasGalerkin    galerkin_;
};

// ------------------------------------------------------------------ PCG
struct PcgConfig
{
    double rel_tol     = 1e-4;
    double abs_tol     = 0.0;
    // > 0: the relative threshold is taken against this initial quantity
    // (r.z or r.r of the zero-guess right-hand side) instead of the solve's own
    double q_ref       = 0.0;
    int    max_iter    = 0;
    bool   use_rz      = true;
    int    check_every = 10;
    // mixed precision: FP32 PCG whose periodic residual recomputation is done
    // in FP64 (with use_rz = false, rel_tol bounds the FP64 true residual)
    bool   mixed = false;
    // mixed precision: stop the FP32 iteration when the tested residual has
    // not halved over this many stop tests (0: the solver's default, 20)
    int    stall_checks = 0;
};

struct PcgStats
{
    int    iterations       = 0;  // PCG iterations (FP32 + FP64 fallback in mixed mode)
    int    outer_iterations = 0;  // 1, or 2 when the mixed mode fell back to FP64
    int    breakdown        = 0;  // 1 non-finite r.z, 2 diverged (mixed mode)
    bool   converged        = false;
    bool   stalled          = false;  // mixed mode: FP32 iteration stopped making progress
    double rz0 = 0, rz = 0;
    double rel_residual = 0;
};

// PCG work vectors of one precision
template <typename V>
struct PcgVecs
{
    DBuf<V> r, z, p[2], Ap, tmp, res, rhs, sol;
    int     n = 0;
    void    ensure(int n_)
    {
        if(n_ == n)
            return;
        for(DBuf<V>* b : {&r, &z, &p[0], &p[1], &Ap, &tmp, &res, &rhs, &sol})
            b->resize(n_);
        n = n_;
    }
};

// block-Jacobi of affine bodies (precond_abd.hip): binv[144 b] = inverse of
// body b's 12x12 diagonal block, gathered from the BSR slots slots[16 b + 4 m + n]
// of its node pairs (m, n); the update applies it to the nodes first + 4 b ..
void launch_body_inverse(int nbody, const int* slots, const double* vals, double* binv,
                         hipStream_t s);
template <typename V>
void launch_update_body(int nbody, int first, V* x, V* r, const V* p, const V* Ap,
                        const double* rz_prev, const double* pap, const V* binv, V* z,
                        double* rz_out, const int* done, hipStream_t s);

enum class Precond
{
    Diag = 0,
    Mas  = 1
};

class Pcg
{
  public:
    Pcg() = default;
    ~Pcg();
    Pcg(const Pcg&)            = delete;
    Pcg& operator=(const Pcg&) = delete;
    // solve A x = b from x = 0 with the reference's iteration structure
    PcgStats solve(const SysView& S, Precond pc, Mas* mas, const double* b, double* x,
                   const PcgConfig& cfg, hipStream_t s);
    // affine bodies: nodes first .. first + 4 nbody - 1 in groups of 4, whose
    // 12x12 diagonal blocks (BSR slots slots16[16 b + 4 m + n]) the block-Jacobi
    // preconditioner inverts whole (ABDPreconditioner) instead of by 3x3 node
    void set_bodies(int first, int nbody, const int* slots16)
    {
        body_first_ = first;
        nbody_      = nbody;
        body_slots_ = slots16;
    }
    // block-Jacobi inverse of the diagonal blocks (DiagPreconditioner, and
    // the affine bodies' 12x12 blocks)
    void build_diag(const Bsr& A, hipStream_t s);
    void apply_diag(int nb, const double* r, double* z, hipStream_t s);
    // FP64 ||b - A x|| / ||b||
    double   true_residual(const SysView& S, const double* b, const double* x, hipStream_t s);
    Profiler* prof = nullptr;

    // Work memory in a device region the caller owns instead of the
    // library's allocations, for FP64 solves of n DOFs: the reduction sets,
    // the control words, the FP64 work vectors and the block-Jacobi inverse
    // are carved from base (workspace_doubles(n) doubles, 256-byte aligned),
    // and the stop flag is read back into host memory held by this object.
    // After binding, no call of this solver allocates; a mixed-precision
    // solve is refused.
    static size_t workspace_doubles(int n);
    void          bind_workspace(double* base, int n);

    // mixed precision: after an FP32 solve that stalls or breaks down, the
    // next 2 mixed solves run in FP64 directly, the pause doubling (up to 32)
    // while the FP32 attempts keep failing; restarted by reset_precision()
    void reset_precision()
    {
        fp32_pause_ = 0;
        fp32_skip_  = 0;
    }

  private:
    // b64/x64 non-null: FP64 residual replacement at every check (mixed mode)
    template <typename V>
    PcgStats solve_t(const SysView& S, Precond pc, Mas* mas, const V* b, V* x,
                     const PcgConfig& cfg, hipStream_t s, const double* b64, double* x64);
    PcgStats solve_mixed(const SysView& S, Precond pc, Mas* mas, const double* b, double* x,
                         const PcgConfig& cfg, hipStream_t s);
    template <typename V>
    PcgVecs<V>& vecs();
    template <typename V>
    const V* dinv() const;
    void ensure(int n);
    // ring of kRing (pAp, rz) slot-set pairs indexed by iteration, plus aux
    // sets for r.r, |b|^2 and the final residual
    static constexpr int kRing = 4;
    double* pap_set(int k) { return red_.data() + (size_t)(2 * (k % kRing)) * kSlotSet; }
    double* rz_set(int k) { return red_.data() + (size_t)(2 * (k % kRing) + 1) * kSlotSet; }
    double* aux_set(int i) { return red_.data() + (size_t)(2 * kRing + i) * kSlotSet; }
    PcgVecs<double> v64_;
    PcgVecs<float>  v32_;
    DBuf<double> red_, dinv_, binv_;
    DBuf<float>  dinv32_, binv32_;
    int          body_first_ = 0, nbody_ = 0;
    const int*   body_slots_ = nullptr;
    template <typename V>
    const V*     binv() const;
    // z = Dinv r over the 3x3 nodes and the bodies, with the optional update
    template <typename V>
    void         apply_block_jacobi(int nb, V* x, V* r, const V* p, const V* Ap,
                                    const double* rz_prev, const double* pap, V* z,
                                    double* rz_out, double* zero0, double* zero1,
                                    const int* done, hipStream_t s);
    DBuf<int>    ctl_;         // [0] stopping check iteration (0: running), [1] breakdown
    DBuf<double> ctl_d_;       // [0] r.z at the last stop test
    int*         h_ctl_ = nullptr;  // pinned: two slots of 4 ints for the stop flag
    hipEvent_t   ev_[2] = {nullptr, nullptr};
    int          n_ = 0;
    // bound to a caller's workspace: the stop flag goes to h_ctl_local_ with a
    // stream synchronisation in place of the pinned copy and its event
    bool         external_ = false;
    int          h_ctl_local_[8] = {};
    int          fp32_pause_ = 0, fp32_skip_ = 0;
};

}  // namespace sgipc
