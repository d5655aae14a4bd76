// tp_cubature.cpp -- host side of the h-adaptive cubature (see tp_cubature.h).
// Algorithm: classed_version/hcubature.cpp:115-1068 and classed_version/converged.h.
#include "tp_cubature.h"

#include <cfloat>
#include <cmath>
#include <cstring>

namespace tpc {

double compute_vol(unsigned dim, const double *hw)
{
    double vol = 1;  // classed_version/hcubature.cpp:134-141
    for (unsigned i = 0; i < dim; ++i) vol *= 2 * hw[i];
    return vol;
}

bool converged(unsigned fdim, const double *val, const double *err, double reqAbs, double reqRel, int norm)
{
    // classed_version/converged.h:6-73
    switch (norm) {
    case INDIVIDUAL:
        for (unsigned j = 0; j < fdim; ++j)
            if (err[j] > reqAbs && err[j] > std::fabs(val[j]) * reqRel) return false;
        return true;
    case PAIRED: {
        unsigned j;
        for (j = 0; j + 1 < fdim; j += 2) {
            double maxerr = err[j] > err[j + 1] ? err[j] : err[j + 1];
            double maxval = val[j] > val[j + 1] ? val[j] : val[j + 1];
            double serr = maxerr > 0 ? 1 / maxerr : 1, sval = maxval > 0 ? 1 / maxval : 1;
            double e0 = err[j] * serr, e1 = err[j + 1] * serr, v0 = val[j] * sval, v1 = val[j + 1] * sval;
            double e = std::sqrt(e0 * e0 + e1 * e1) * maxerr;
            double v = std::sqrt(v0 * v0 + v1 * v1) * maxval;
            if (e > reqAbs && e > v * reqRel) return false;
        }
        if (j < fdim && err[j] > reqAbs && err[j] > std::fabs(val[j]) * reqRel) return false;
        return true;
    }
    case L1: {
        double e = 0, v = 0;
        for (unsigned j = 0; j < fdim; ++j) {
            e += err[j];
            v += std::fabs(val[j]);
        }
        return e <= reqAbs || e <= v * reqRel;
    }
    case LINF: {
        double e = 0, v = 0;
        for (unsigned j = 0; j < fdim; ++j) {
            if (err[j] > e) e = err[j];
            if (std::fabs(val[j]) > v) v = std::fabs(val[j]);
        }
        return e <= reqAbs || e <= v * reqRel;
    }
    case L2: {
        double maxerr = 0, maxval = 0, e = 0, v = 0;
        for (unsigned j = 0; j < fdim; ++j) {
            if (err[j] > maxerr) maxerr = err[j];
            if (std::fabs(val[j]) > maxval) maxval = std::fabs(val[j]);
        }
        double serr = maxerr > 0 ? 1 / maxerr : 1, sval = maxval > 0 ? 1 / maxval : 1;
        for (unsigned j = 0; j < fdim; ++j) {
            double a = err[j] * serr, b = std::fabs(val[j]) * sval;
            e += a * a;
            v += b * b;
        }
        e = std::sqrt(e) * maxerr;
        v = std::sqrt(v) * maxval;
        return e <= reqAbs || e <= v * reqRel;
    }
    }
    return true;
}

unsigned rule_points(unsigned dim)
{
    if (dim == 1) return 15;
    return 1 + 4 * dim + 2 * dim * (dim - 1) + (1u << dim);
}

// Genz-Malik generators (classed_version/hcubature.cpp:456-503)
static const double kL2 = 0.3585685828003180919906451539079374954541;  // sqrt(9/70)
static const double kL4 = 0.9486832980505137995996680633298155601160;  // sqrt(9/10)
static const double kL5 = 0.6882472016116852977216287342936235251269;  // sqrt(9/19)

// Kronrod-15 abscissae / weights and Gauss-7 weights (QUADPACK, Piessens et al. 1983)
static const double kXgk[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                               0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                               0.586087235467691130294144838258730, 0.405845151377397166906606412076961,
                               0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
static const double kWg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                              0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
static const double kWgk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                               0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                               0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                               0.204432940075298892414161999234649, 0.209482141084727828012999174891714};

template <unsigned MD, unsigned MF>
void rule_points_of(unsigned dim, const RegionT<MD, MF> &R, double *pts)
{
    const double *c = R.c, *h = R.hw;
    if (dim == 1) {  // classed_version/hcubature.cpp:663-681
        double *P = pts;
        *P++ = c[0];
        for (int j = 0; j < 3; ++j) {
            double w = h[0] * kXgk[2 * j + 1];
            *P++ = c[0] - w;
            *P++ = c[0] + w;
        }
        for (int j = 0; j < 4; ++j) {
            double w = h[0] * kXgk[2 * j];
            *P++ = c[0] - w;
            *P++ = c[0] + w;
        }
        return;
    }
    double p[kMaxDim], w2[kMaxDim], w4[kMaxDim], w5[kMaxDim];
    for (unsigned i = 0; i < dim; ++i) {
        p[i] = c[i];
        w2[i] = h[i] * kL2;
        w4[i] = h[i] * kL4;
        w5[i] = h[i] * kL5;
    }
    double *P = pts;
    auto emit = [&]() {
        std::memcpy(P, p, sizeof(double) * dim);
        P += dim;
    };
    emit();  // centre
    for (unsigned i = 0; i < dim; ++i) {
        p[i] = c[i] - w2[i]; emit();
        p[i] = c[i] + w2[i]; emit();
        p[i] = c[i] - w4[i]; emit();
        p[i] = c[i] + w4[i]; emit();
        p[i] = c[i];
    }
    for (unsigned i = 0; i + 1 < dim; ++i) {  // (lambda4, lambda4) pairs
        p[i] = c[i] - w4[i];
        for (unsigned j = i + 1; j < dim; ++j) {
            p[j] = c[j] - w4[j]; emit();
            p[i] = c[i] + w4[i]; emit();
            p[j] = c[j] + w4[j]; emit();
            p[i] = c[i] - w4[i]; emit();
            p[j] = c[j];
        }
        p[i] = c[i];
    }
    unsigned signs = 0;  // 2^dim corners, Gray-code order starting from all '+'
    for (unsigned i = 0; i < dim; ++i) p[i] = c[i] + w5[i];
    for (unsigned i = 0;; ++i) {
        emit();
        unsigned d = (unsigned)__builtin_ctz(~i);
        if (d >= dim) break;
        signs ^= 1u << d;
        p[d] = (signs & (1u << d)) ? c[d] - w5[d] : c[d] + w5[d];
    }
}

template <unsigned MD, unsigned MF>
void rule_apply(unsigned dim, unsigned fdim, RegionT<MD, MF> &R, const double *vals)
{
    if (dim == 1) {  // GK15 with QUADPACK's error rescaling, classed_version/hcubature.cpp:688-754
        for (unsigned k = 0; k < fdim; ++k) {
            const double *vk = vals + k;
            double hw = R.hw[0];
            double rg = vk[0] * kWg[3], rk = vk[0] * kWgk[7], ra = std::fabs(rk);
            size_t n = 1;
            for (int j = 0; j < 3; ++j) {
                int j2 = 2 * j + 1;
                double v = vk[fdim * n] + vk[fdim * n + fdim];
                rg += kWg[j] * v;
                rk += kWgk[j2] * v;
                ra += kWgk[j2] * (std::fabs(vk[fdim * n]) + std::fabs(vk[fdim * n + fdim]));
                n += 2;
            }
            for (int j = 0; j < 4; ++j) {
                int j2 = 2 * j;
                rk += kWgk[j2] * (vk[fdim * n] + vk[fdim * n + fdim]);
                ra += kWgk[j2] * (std::fabs(vk[fdim * n]) + std::fabs(vk[fdim * n + fdim]));
                n += 2;
            }
            R.val[k] = rk * hw;
            double mean = rk * 0.5, rasc = kWgk[7] * std::fabs(vk[0] - mean);
            n = 1;
            for (int j = 0; j < 3; ++j) {
                int j2 = 2 * j + 1;
                rasc += kWgk[j2] * (std::fabs(vk[fdim * n] - mean) + std::fabs(vk[fdim * n + fdim] - mean));
                n += 2;
            }
            for (int j = 0; j < 4; ++j) {
                int j2 = 2 * j;
                rasc += kWgk[j2] * (std::fabs(vk[fdim * n] - mean) + std::fabs(vk[fdim * n + fdim] - mean));
                n += 2;
            }
            double err = std::fabs(rk - rg) * hw;
            ra *= hw;
            rasc *= hw;
            if (rasc != 0 && err != 0) {
                double scale = std::pow((200 * err / rasc), 1.5);
                err = (scale < 1) ? rasc * scale : rasc;
            }
            if (ra > DBL_MIN / (50 * DBL_EPSILON)) {
                double min_err = 50 * DBL_EPSILON * ra;
                if (min_err > err) err = min_err;
            }
            R.err[k] = err;
        }
        R.split = 0;
    } else {  // Genz-Malik 7/5, classed_version/hcubature.cpp:515-575
        const int id = (int)dim;
        const double w1 = (double)(12824 - 9120 * id + 400 * id * id) / 19683.0;
        const double w3 = (double)(1820 - 400 * id) / 19683.0;
        const double w5 = 6859.0 / 19683.0 / (double)(1u << dim);
        const double wE1 = (double)(729 - 950 * id + 50 * id * id) / 729.0;
        const double wE3 = (double)(265 - 100 * id) / 1458.0;
        const double wt2 = 980. / 6561., wt4 = 200. / 19683., wtE2 = 245. / 486., wtE4 = 25. / 729.;
        const double ratio = (kL2 * kL2) / (kL4 * kL4);
        double diff[kMaxDim] = {0};
        for (unsigned j = 0; j < fdim; ++j) {
            const double *v = vals + j;
            double val0 = v[0], sum2 = 0, sum3 = 0, sum4 = 0, sum5 = 0;
            unsigned k0 = 1, k;
            for (k = 0; k < dim; ++k) {
                double v0 = v[fdim * (k0 + 4 * k)], v1 = v[fdim * (k0 + 4 * k + 1)];
                double v2 = v[fdim * (k0 + 4 * k + 2)], v3 = v[fdim * (k0 + 4 * k + 3)];
                sum2 += v0 + v1;
                sum3 += v2 + v3;
                diff[k] += std::fabs(v0 + v1 - 2 * val0 - ratio * (v2 + v3 - 2 * val0));
            }
            k0 += 4 * k;
            for (k = 0; k < 2 * dim * (dim - 1); ++k) sum4 += v[fdim * (k0 + k)];
            k0 += k;
            for (k = 0; k < (1u << dim); ++k) sum5 += v[fdim * (k0 + k)];
            double result = R.vol * (w1 * val0 + wt2 * sum2 + w3 * sum3 + wt4 * sum4 + w5 * sum5);
            double res5th = R.vol * (wE1 * val0 + wtE2 * sum2 + wE3 * sum3 + wtE4 * sum4);
            R.val[j] = result;
            R.err[j] = std::fabs(res5th - result);
        }
        double maxdiff = 0;
        unsigned dd = 0;
        for (unsigned i = 0; i < dim; ++i)
            if (diff[i] > maxdiff) {
                maxdiff = diff[i];
                dd = i;
            }
        R.split = dd;
    }
    double em = 0;
    for (unsigned k = 0; k < fdim; ++k)
        if (R.err[k] > em) em = R.err[k];
    R.errmax = em;
    R.corr = 0.0;
    R.cvar = 0.0;
}

template <unsigned MD, unsigned MF>
bool AdaptiveIntegralT<MD, MF>::init(unsigned dim, unsigned fdim, const double *xmin, const double *xmax, size_t maxEval,
                            double reqAbs, double reqRel, int norm)
{
    if (dim == 0 || dim > MD || fdim == 0 || fdim > MF) return false;
    if (fdim <= 1) norm = INDIVIDUAL;  // classed_version/hcubature.cpp:919-920
    if (norm < 0 || norm > LINF) return false;
    dim_ = dim;
    fdim_ = fdim;
    maxEval_ = maxEval;
    reqAbs_ = reqAbs;
    reqRel_ = reqRel;
    norm_ = norm;
    npts_ = rule_points(dim);
    pool_.clear();
    heap_.clear();
    pending_.clear();
    for (unsigned k = 0; k < MF; ++k) sval_[k] = serr_[k] = 0;
    RegionType R{};
    for (unsigned i = 0; i < dim; ++i) {  // classed_version/hcubature.cpp:160-172
        R.c[i] = 0.5 * (xmin[i] + xmax[i]);
        R.hw[i] = 0.5 * (xmax[i] - xmin[i]);
    }
    R.vol = compute_vol(dim, R.hw);
    R.errmax = HUGE_VAL;
    R.corr = R.cvar = 0.0;
    pool_.push_back(R);
    pending_.push_back(0);
    numEval_ = npts_;
    rounds_ = 0;
    done_ = converged_ = false;
    return true;
}

template <unsigned MD, unsigned MF>
void AdaptiveIntegralT<MD, MF>::heap_push(uint32_t idx)
{
    // classed_version/hcubature.cpp:808-835
    const RegionType &r = pool_[idx];
    for (unsigned i = 0; i < fdim_; ++i) {
        sval_[i] += r.val[i];
        serr_[i] += r.err[i];
    }
    size_t ins = heap_.size();
    heap_.push_back(HeapEnt{r.errmax, idx});
    const double key = r.errmax;
    HeapEnt *H = heap_.data();
    while (ins) {
        size_t parent = (ins - 1) / 2;
        if (key <= H[parent].key) break;
        H[ins] = H[parent];
        ins = parent;
    }
    H[ins] = HeapEnt{key, idx};
}

template <unsigned MD, unsigned MF>
uint32_t AdaptiveIntegralT<MD, MF>::heap_pop()
{
    // classed_version/hcubature.cpp:853-890: the last item moves to the root and sinks while a
    // child's key is larger (the left child wins unless the right one is strictly larger).
    // Written with a hole and conditional moves; the resulting arrangement is the same.
    uint32_t ret = heap_[0].idx;
    const size_t n = heap_.size() - 1;
    const HeapEnt moved = heap_[n];
    heap_.pop_back();
    if (n) {
        HeapEnt *H = heap_.data();
        const double km = moved.key;
        size_t i = 0, child;
        while ((child = 2 * i + 1) < n) {
            __builtin_prefetch(H + 4 * i + 3);
            const double kc = H[child].key;
            size_t largest = (kc <= km) ? i : child;
            double kl = (kc <= km) ? km : kc;
            const size_t c2 = child + 1;
            if (c2 < n) {
                const double k2 = H[c2].key;
                largest = (kl < k2) ? c2 : largest;
            }
            if (largest == i) break;
            H[i] = H[largest];
            i = largest;
        }
        H[i] = moved;
    }
    const RegionType &r = pool_[ret];
    for (unsigned k = 0; k < fdim_; ++k) {
        sval_[k] -= r.val[k];
        serr_[k] -= r.err[k];
    }
    return ret;
}

template <unsigned MD, unsigned MF>
void AdaptiveIntegralT<MD, MF>::absorb()
{
    for (uint32_t idx : pending_) heap_push(idx);
    pending_.clear();
}

template <unsigned MD, unsigned MF>
bool AdaptiveIntegralT<MD, MF>::advance()
{
    // one iteration of the main loop, classed_version/hcubature.cpp:938-989
    if (done_) return true;
    if (!(numEval_ < maxEval_ || !maxEval_)) {
        done_ = true;
        return true;
    }
    if (converged(fdim_, sval_, serr_, reqAbs_, reqRel_, norm_)) {
        done_ = converged_ = true;
        return true;
    }
    double eerr[MF], eval_[MF];
    for (unsigned j = 0; j < fdim_; ++j) {
        eerr[j] = serr_[j];
        eval_[j] = sval_[j];
    }
    do {
        uint32_t a = heap_pop();
        for (unsigned j = 0; j < fdim_; ++j) eerr[j] -= pool_[a].err[j];
        // cut_region (classed_version/hcubature.cpp:206-218): the kept half halves its
        // volume, the new half recomputes it from the widths.
        RegionType &A = pool_[a];
        unsigned d = A.split;
        A.hw[d] *= 0.5;
        A.vol *= 0.5;
        RegionType B = A;
        B.vol = compute_vol(dim_, B.hw);
        A.c[d] -= A.hw[d];
        B.c[d] += A.hw[d];
        pool_.push_back(B);  // invalidates A
        uint32_t b = (uint32_t)(pool_.size() - 1);
        pending_.push_back(a);
        pending_.push_back(b);
        numEval_ += npts_ * 2;
        if (converged(fdim_, eval_, eerr, reqAbs_, reqRel_, norm_)) break;
    } while (!heap_.empty() && (numEval_ < maxEval_ || !maxEval_));
    ++rounds_;
    return false;
}

template <unsigned MD, unsigned MF>
void AdaptiveIntegralT<MD, MF>::result(double *val, double *err) const
{
    // final re-sum in heap order, classed_version/hcubature.cpp:1000-1008; the leaf corrections are
    // summed in the same pass (each sum keeps its own order, so the bits equal separate passes)
    for (unsigned j = 0; j < fdim_; ++j) val[j] = err[j] = 0;
    double c = 0;
    for (const HeapEnt &h : heap_) {
        const RegionType &r = pool_[h.idx];
        for (unsigned j = 0; j < fdim_; ++j) {
            val[j] += r.val[j];
            err[j] += r.err[j];
        }
        c += r.corr;
    }
    if (c != 0.0) val[0] += c;
}

template <unsigned MD, unsigned MF>
double AdaptiveIntegralT<MD, MF>::correction_var() const
{
    double v = 0;
    for (const HeapEnt &h : heap_) v += pool_[h.idx].cvar;
    return v;
}

template <unsigned MD, unsigned MF>
double AdaptiveIntegralT<MD, MF>::correction() const
{
    double c = 0;
    for (const HeapEnt &h : heap_) c += pool_[h.idx].corr;
    return c;
}

template class AdaptiveIntegralT<kMaxDim, kMaxFdim>;
template class AdaptiveIntegralT<3, 1>;
template class AdaptiveIntegralT<3, 2>;
template class AdaptiveIntegralT<1, 1>;
template class AdaptiveIntegralT<2, 1>;
template class AdaptiveIntegralT<4, 1>;
template class AdaptiveIntegralT<6, 1>;
template void rule_points_of<kMaxDim, kMaxFdim>(unsigned, const RegionT<kMaxDim, kMaxFdim> &, double *);
template void rule_points_of<3, 1>(unsigned, const RegionT<3, 1> &, double *);
template void rule_points_of<3, 2>(unsigned, const RegionT<3, 2> &, double *);
template void rule_points_of<1, 1>(unsigned, const RegionT<1, 1> &, double *);
template void rule_points_of<2, 1>(unsigned, const RegionT<2, 1> &, double *);
template void rule_points_of<4, 1>(unsigned, const RegionT<4, 1> &, double *);
template void rule_points_of<6, 1>(unsigned, const RegionT<6, 1> &, double *);
template void rule_apply<kMaxDim, kMaxFdim>(unsigned, unsigned, RegionT<kMaxDim, kMaxFdim> &, const double *);
template void rule_apply<3, 1>(unsigned, unsigned, RegionT<3, 1> &, const double *);
template void rule_apply<3, 2>(unsigned, unsigned, RegionT<3, 2> &, const double *);
template void rule_apply<1, 1>(unsigned, unsigned, RegionT<1, 1> &, const double *);
template void rule_apply<2, 1>(unsigned, unsigned, RegionT<2, 1> &, const double *);
template void rule_apply<4, 1>(unsigned, unsigned, RegionT<4, 1> &, const double *);
template void rule_apply<6, 1>(unsigned, unsigned, RegionT<6, 1> &, const double *);

}  // namespace tpc
