// tp_cubature.h -- h-adaptive cubature state machine (host C++).
//
// The algorithm is the one of classed_version/hcubature.cpp (Genz & Malik 1980
// degree-7/5 rule, Berntsen-Espelid-Genz vector integrands, Gladwell's batch
// selection): a max-heap of regions keyed on their largest error, running sums
// of value and error updated on every push and pop, convergence tested with the
// error norms of classed_version/converged.h.  This implementation separates
// the decision logic (advance) from region evaluation, so a round's regions can
// be evaluated anywhere: by a host callback (tp_hcubature_v) or, for many
// integrals at once, by one GPU launch (tp_map3_batch).  The decision sequence
// of each integral is that of the one-at-a-time algorithm.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace tpc {

constexpr unsigned kMaxDim = 8;
constexpr unsigned kMaxFdim = 4;

// One region: centre, half-widths, volume, split dimension and rule results.  The
// array sizes are compile-time bounds on dim and fdim; the batched engines use
// exact-size regions (3-D, 1 or 2 components) so the region pool stays compact.
template <unsigned MD, unsigned MF>
struct RegionT {
    double c[MD], hw[MD];
    double vol;
    unsigned split;
    double val[MF], err[MF];
    double errmax;
    double corr;  // additive correction of val[0] carried by a leaf (surrogate checks), 0 otherwise
    double cvar;  // variance estimate of corr
};
typedef RegionT<kMaxDim, kMaxFdim> Region;

enum Norm { INDIVIDUAL = 0, PAIRED = 1, L2 = 2, L1 = 3, LINF = 4 };

bool converged(unsigned fdim, const double *val, const double *err, double reqAbs, double reqRel, int norm);
double compute_vol(unsigned dim, const double *hw);

template <unsigned MD, unsigned MF>
class AdaptiveIntegralT {
  public:
    typedef RegionT<MD, MF> RegionType;
    // Sets up the first region (pending evaluation).  Returns false on bad arguments.
    bool init(unsigned dim, unsigned fdim, const double *xmin, const double *xmax, size_t maxEval, double reqAbs,
              double reqRel, int norm);
    // Regions awaiting evaluation (indices into regions()).
    const std::vector<uint32_t> &pending() const { return pending_; }
    std::vector<RegionType> &regions() { return pool_; }
    const std::vector<RegionType> &regions() const { return pool_; }
    // After the pending regions were evaluated (val, err, split filled): push them.
    void absorb();
    // Convergence test + Gladwell selection of the next round.  Returns true when
    // the integral is finished (then result() is valid and pending() is empty).
    bool advance();
    void result(double *val, double *err) const;
    // sum of the leaf corrections (part of result()'s val[0])
    double correction() const;
    double correction_var() const;
    size_t num_eval() const { return numEval_; }
    size_t num_regions() const { return pool_.size(); }
    unsigned dim() const { return dim_; }
    unsigned fdim() const { return fdim_; }
    long long rounds() const { return rounds_; }
    bool converged_flag() const { return converged_; }
    size_t heap_size() const { return heap_.size(); }
    // host bytes held by the region pool, heap and pending list (capacity, kept across init())
    size_t bytes_reserved() const
    {
        return pool_.capacity() * sizeof(RegionType) + heap_.capacity() * sizeof(HeapEnt) +
               pending_.capacity() * sizeof(uint32_t);
    }

  private:
    void heap_push(uint32_t idx);
    uint32_t heap_pop();
    unsigned dim_ = 0, fdim_ = 0;
    size_t maxEval_ = 0, numEval_ = 0, npts_ = 0;
    double reqAbs_ = 0, reqRel_ = 0;
    int norm_ = L1;
    std::vector<RegionType> pool_;
    struct HeapEnt {
        double key;    // errmax of the region (kept here so sifts touch only the heap array)
        uint32_t idx;  // index into pool_
    };
    std::vector<HeapEnt> heap_;  // binary max-heap on errmax
    std::vector<uint32_t> pending_;
    double sval_[MF] = {0}, serr_[MF] = {0};
    long long rounds_ = 0;
    bool done_ = false;
    bool converged_ = false;  // finished by the tolerance test (not by maxEval)
};
typedef AdaptiveIntegralT<kMaxDim, kMaxFdim> AdaptiveIntegral;  // any dim <= 8, fdim <= 4
typedef AdaptiveIntegralT<3, 1> AdaptiveIntegral31;             // <Map^3>
typedef AdaptiveIntegralT<3, 2> AdaptiveIntegral32;             // Gamma_i (re, im)

// Number of Genz-Malik points of a dim-dimensional region (dim >= 2), 15 for dim 1.
unsigned rule_points(unsigned dim);
// Writes the rule's points of region R (in the reference's order) to pts (npts*dim).
template <unsigned MD, unsigned MF>
void rule_points_of(unsigned dim, const RegionT<MD, MF> &R, double *pts);
// Fills R.val/err/split/errmax from its point values v (npts*fdim).
template <unsigned MD, unsigned MF>
void rule_apply(unsigned dim, unsigned fdim, RegionT<MD, MF> &R, const double *v);

}  // namespace tpc
