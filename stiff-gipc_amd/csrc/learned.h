// learned.h -- the learned coarse-level warm start (learned.hip).
#pragma once

#include <stdint.h>

#include "linsys.h"

namespace sgipc
{

// network shape and the packed parameter layout (FP32 master copy; the BF16
// copy of the matrices uses the same offsets)
constexpr int kIn = 24, kHid = 128, kOut = 12;
constexpr int kOffW1  = 0;                       // [kIn][kHid]
constexpr int kOffB1  = kOffW1 + kIn * kHid;     // [kHid]
constexpr int kOffW2  = kOffB1 + kHid;           // [kHid][kHid]
constexpr int kOffB2  = kOffW2 + kHid * kHid;    // [kHid]
constexpr int kOffW3  = kOffB2 + kHid;           // [kHid][kOut]
constexpr int kOffB3  = kOffW3 + kHid * kOut;    // [kOut]
constexpr int kParams = kOffB3 + kOut;

struct MlpW
{
    const uint16_t *W1, *W2, *W3;  // BF16 bits
    const float *   b1, *b2, *b3;
};

class Mlp
{
  public:
    void   init(uint64_t seed, hipStream_t s);
    void   set(const float* host, hipStream_t s);
    void   get(float* host, hipStream_t s) const;
    // full-batch Adam for `epochs` steps; returns the final loss
    double train(const float* X, const float* T, int n, int epochs, double lr, hipStream_t s);
    // MSE loss and its gradient (host, kParams floats) at the current weights
    double loss_grad(const float* X, const float* T, int n, float* grad_host, hipStream_t s);
    void   infer(const float* feat_dev, int n, float* y_dev, hipStream_t s) const;
    bool   ready() const { return ready_; }

  private:
    MlpW   view() const;
    void   ensure_batch(int n);
    double pass(int n, hipStream_t s);
    DBuf<float>    w_, m_, v_, g_;
    DBuf<uint16_t> wb_;
    DBuf<float>    X_, T_, Y_, dY_, loss_, H1_, H2_, dH1_, dH2_;
    DBuf<double>   lsum_;
    int            cap_ = 0, step_ = 0;
    bool           ready_ = false;
};

// affine coarse spaces on the coarsest-level aggregates of the MAS hierarchy
class Coarse
{
  public:
    bool setup(const Mas& mas, hipStream_t s);  // false: no coarse level
    // aggregates (free FEM nodes), centroids, Galerkin 12x12 systems, q_c and
    // the network features for right-hand side b
    void build(int nn, int nfem, const int* btype, const double* x, const Bsr& A, const Coo* coo,
               const double* b, hipStream_t s);
    // x0 = P q_c, corrected by the network when given
    void guess(const Mlp* net, double* x0, hipStream_t s);
    // (features, target) of every aggregate with a defined scale; targets
    // from the affine least-squares fit of the solution sol
    int  samples(const double* sol, float* feat, float* target, int cap, hipStream_t s);
    int  count() const { return nc_; }

  private:
    void ensure(int nc, int nn);
    int           L_ = 0, lev_ = 0, nc_ = 0, off_ = 0, cap_ = 0, ncap_ = 0, nn_ = 0;
    static constexpr int kMaxAggregateNodes = 4096;
    const int*    chain_ = nullptr;
    const double* x_     = nullptr;
    DBuf<int>      agg_, nodes_, ptr_, okf_;
    DBuf<unsigned> cnt_, cur_;
    DBuf<double>   cen_, Ac_, rc_, qc_, qfit_, scale_;
    DBuf<float>    feat_, y_;
};

}  // namespace sgipc
