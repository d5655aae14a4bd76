// tp_census.cpp -- algorithmic operation census of one <Map^3> integrand point.
//
// The kernel's own per-point and per-node functions (tp_device.h) are instantiated
// on the host with a counting scalar: every +, -, *, / is one flop (a fused
// multiply-add is counted as its multiply and its add, i.e. two), every exp / log
// is one elementary-function call.  Comparisons, selects and loads are not counted.
#include <cmath>

#include "tp_power_dev.h"

namespace {
struct OpCounts {
    long long flops = 0, trans = 0;
};
thread_local OpCounts g_cnt;

struct Cnt {
    double v;
    Cnt(double x = 0) : v(x) {}
};
inline Cnt operator+(Cnt a, Cnt b) { ++g_cnt.flops; return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { ++g_cnt.flops; return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { ++g_cnt.flops; return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { ++g_cnt.flops; return Cnt(a.v / b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline Cnt &operator+=(Cnt &a, Cnt b) { a = a + b; return a; }
inline Cnt &operator*=(Cnt &a, Cnt b) { a = a * b; return a; }
inline bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
// elementary functions of the power-spectrum point (found by argument-dependent lookup)
inline Cnt pow(Cnt a, Cnt b) { ++g_cnt.trans; return Cnt(std::pow(a.v, b.v)); }
inline Cnt exp(Cnt a) { ++g_cnt.trans; return Cnt(std::exp(a.v)); }
inline Cnt log(Cnt a) { ++g_cnt.trans; return Cnt(std::log(a.v)); }
inline Cnt sqrt(Cnt a) { ++g_cnt.trans; return Cnt(std::sqrt(a.v)); }
inline Cnt log1p(Cnt a) { ++g_cnt.trans; return Cnt(std::log1p(a.v)); }
inline Cnt fabs(Cnt a) { return Cnt(std::fabs(a.v)); }
inline double tp_dbl(Cnt a) { return a.v; }
}  // namespace

template <> struct TpMath<Cnt> {
    static Cnt ex(Cnt x) { ++g_cnt.trans; return Cnt(std::exp(x.v)); }
    static Cnt lg(Cnt x) { ++g_cnt.trans; return Cnt(std::log(x.v)); }
    static Cnt rcp(Cnt x) { ++g_cnt.flops; return Cnt(1.0 / x.v); }
};

// out: [0] flops per node, [1] elementary calls per node, [2] flops per point outside
// the node loop, [3] elementary calls per point outside the node loop, [4] nodes per point
extern "C" int tp_census_map3_point(double *out)
{
    TpNodeT<Cnt> N;
    double *dst = nullptr;
    (void)dst;
    // representative node values (the census does not depend on them except through branches)
    Cnt vals[sizeof(TpNodeT<Cnt>) / sizeof(Cnt)];
    for (auto &v : vals) v = Cnt(0.5);
    N = *reinterpret_cast<TpNodeT<Cnt> *>(vals);
    N.acap = Cnt(10.0);
    TpEHT<Cnt> E;
    Cnt ev[sizeof(TpEHT<Cnt>) / sizeof(Cnt)];
    for (auto &v : ev) v = Cnt(0.7);
    E = *reinterpret_cast<TpEHT<Cnt> *>(ev);
    E.tab = TpPkTable{};  // the census counts the Eisenstein & Hu form
    // per point: l3, the hoisted scale-free quantities, the three filters and the final product
    g_cnt = OpCounts();
    Cnt l1(300.), l2(500.), phi(1.0);
    Cnt s = l1 * l1 + l2 * l2 + Cnt(2) * l1 * l2 * TpMath<Cnt>::ex(phi);  // cos ~ one elementary call
    Cnt l3 = TpMath<Cnt>::ex(s);                                           // sqrt ~ one elementary call
    (void)l3;
    TpPointPreT<Cnt> P;
    tp_point_pre<Cnt>(Cnt(300.), Cnt(500.), Cnt(600.), &P);
    Cnt f = l1 * l2 * Cnt(1.0) * tp_uhat<Cnt>(l1 * Cnt(1e-3)) * tp_uhat<Cnt>(l2 * Cnt(1e-3)) *
            tp_uhat<Cnt>(Cnt(600.) * Cnt(1e-3));
    (void)f;
    OpCounts point = g_cnt;
    // per node: wavenumbers, logs, bispectrum, weight and accumulation (tp_bkappa_sum body)
    g_cnt = OpCounts();
    Cnt L[3] = {Cnt(300.), Cnt(500.), Cnt(600.)};
    Cnt q[3] = {L[0] * N.ks, L[1] * N.ks, L[2] * N.ks};
    Cnt lq[3] = {P.ll[0] + N.lnks, P.ll[1] + N.lnks, P.ll[2] + N.lnks};
    Cnt b = tp_bispec_node<Cnt>(N, E, q, lq, P.f2t, P.lr1, P.r2);
    Cnt acc(0.0);
    acc += N.wz * b;
    OpCounts node = g_cnt;
    out[0] = (double)node.flops;
    out[1] = (double)node.trans;
    out[2] = (double)point.flops;
    out[3] = (double)point.trans;
    out[4] = (double)TP_NNODES;
    return 0;
}

// out: [0] flops and [1] elementary calls of one P_E term of BiHalofit (tp_pse_node, a function of one
// wavenumber; a node evaluation of tp_bispec_node makes three)
extern "C" int tp_census_map3_pse(double *out)
{
    Cnt vals[sizeof(TpNodeT<Cnt>) / sizeof(Cnt)];
    for (auto &v : vals) v = Cnt(0.5);
    TpNodeT<Cnt> N = *reinterpret_cast<TpNodeT<Cnt> *>(vals);
    N.acap = Cnt(10.0);
    Cnt ev[sizeof(TpEHT<Cnt>) / sizeof(Cnt)];
    for (auto &v : ev) v = Cnt(0.7);
    TpEHT<Cnt> E = *reinterpret_cast<TpEHT<Cnt> *>(ev);
    E.tab = TpPkTable{};
    g_cnt = OpCounts();
    Cnt pse = tp_pse_node<Cnt>(N, E, Cnt(0.4), Cnt(-0.9));
    (void)pse;
    out[0] = (double)g_cnt.flops;
    out[1] = (double)g_cnt.trans;
    return 0;
}

// out: [0] flops per Ogata term, [1] elementary calls per term (a Gamma cubature point has
// 3 (prec_k - 1) terms), [2] flops and [3] elementary calls of the per-point shape (including the
// node-and-shape part of the one-halo term, formed once per point since all its terms share both)
extern "C" int tp_census_gamma_term(double *out)
{
    Cnt vals[sizeof(TpNodeT<Cnt>) / sizeof(Cnt)];
    for (auto &v : vals) v = Cnt(0.5);
    TpNodeT<Cnt> N = *reinterpret_cast<TpNodeT<Cnt> *>(vals);
    N.acap = Cnt(10.0);
    Cnt ev[sizeof(TpEHT<Cnt>) / sizeof(Cnt)];
    for (auto &v : ev) v = Cnt(0.7);
    TpEHT<Cnt> E = *reinterpret_cast<TpEHT<Cnt> *>(ev);
    E.tab = TpPkTable{};
    g_cnt = OpCounts();
    // shape: cos psi, sin psi, cos phi (3 calls), u3^2, sqrt (1 call), ln A_j (3 calls)
    Cnt cps = TpMath<Cnt>::ex(Cnt(0.3)), sps = TpMath<Cnt>::ex(Cnt(0.3)), cph = TpMath<Cnt>::ex(Cnt(0.3));
    Cnt u3 = TpMath<Cnt>::ex(cps * cps + sps * sps + Cnt(2) * cps * sps * cph);
    TpPointPreT<Cnt> P;
    tp_point_pre<Cnt>(cps, sps, u3, &P);
    for (int j = 0; j < 3; ++j) (void)TpMath<Cnt>::lg(Cnt(2.0));
    TpShapeNodeT<Cnt> S;
    tp_shape_node<Cnt>(N, P.lr1, P.r2, &S);
    OpCounts shape = g_cnt;
    g_cnt = OpCounts();
    Cnt acc(0.0);
    acc += tp_ogata_term<Cnt>(N, E, P, S, Cnt(30.0), Cnt(3.4), Cnt(0.1));
    OpCounts term = g_cnt;
    out[0] = (double)term.flops;
    out[1] = (double)term.trans;
    out[2] = (double)shape.flops;
    out[3] = (double)shape.trans;
    return 0;
}

// out: [0] flops and [1] elementary calls (pow, exp, log, sqrt) of one <Map^2> integrand point
// (Halofit, no T17 factor, no shape noise); [2], [3] the same for one P_kappa Limber point; [4], [5]
// for one precomputed z-node of a fixed-node Limber sum (the k-dependent part only, as Pell and T1
// evaluate it); divisions count as flops
extern "C" int tp_census_power_point(double *out)
{
    TpCosmoConst C{};
    C.h = 0.7; C.sigma8 = 0.82; C.omb = 0.047; C.omc = 0.253; C.ns = 0.97; C.w = -1; C.om = 0.3; C.ow = 0.7;
    C.zmax = 3; C.norm = 1; C.eh_fb = 0.16; C.eh_sqrt_alnu = 0.9; C.eh_sh = 100; C.eh_omh2 = 0.15;
    C.eh_theta2 = 1.02; C.eh_cL = 2.0; C.limber_pref = 1e-9; C.n_z = 4;
    C.tab = TpPkTable{};  // the census counts the Eisenstein & Hu form
    double tab[6 * 4];
    for (int i = 0; i < 24; ++i) tab[i] = 1.0 + 0.1 * (i % 4);
    g_cnt = OpCounts();
    Cnt m = tp_map2_point<Cnt>(C, tab, 4, Cnt(2000.), Cnt(0.7), 5e-4, 0, 0.0, 1.0);
    (void)m;
    OpCounts map2 = g_cnt;
    g_cnt = OpCounts();
    Cnt p = tp_limber_pk<Cnt>(C, tab, 4, Cnt(2000.), Cnt(0.7), 0);
    (void)p;
    OpCounts pk = g_cnt;
    // one z-node of a fixed-node Limber sum: k = ell / f_K, the k-dependent Halofit part, pref * P
    g_cnt = OpCounts();
    TpHaloZT<Cnt> h;
    h.f1 = Cnt(1.0); h.f2 = Cnt(0.9); h.f3 = Cnt(1.1); h.a = Cnt(2.0); h.b = Cnt(0.5); h.c = Cnt(0.7);
    h.gam = Cnt(0.3); h.alpha = Cnt(1.5); h.beta = Cnt(0.4); h.xnu = Cnt(1e5); h.lf3c = Cnt(0.5);
    Cnt k = Cnt(2000.) / Cnt(1500.), lk = Cnt(7.6) - Cnt(7.3);
    Cnt node = Cnt(1.0) * Cnt(2.0) * tp_halofit_k<Cnt>(C, Cnt(0.8), Cnt(2.5), Cnt(0.9), h, k, lk);
    (void)node;
    OpCounts nd = g_cnt;
    out[0] = (double)map2.flops;
    out[1] = (double)map2.trans;
    out[2] = (double)pk.flops;
    out[3] = (double)pk.trans;
    out[4] = (double)nd.flops;
    out[5] = (double)nd.trans;
    return 0;
}
