// tp_device.h -- device functions shared by the integrand, region, Gamma and
// surrogate kernels.  See tp_map3.hip for the overview.
#pragma once
#include <type_traits>

#include "tp_internal.h"

namespace tpd {
constexpr double kPi = 3.14159265358979323846;
}

template <class T> struct TpMath;
template <> struct TpMath<double> {
    static __host__ __device__ __forceinline__ double ex(double x) { return exp(x); }
    static __host__ __device__ __forceinline__ double lg(double x) { return log(x); }
    static __host__ __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }  // IEEE division
};
template <> struct TpMath<float> {
    // native transcendental unit: v_exp_f32 / v_log_f32 with a scale, v_rcp_f32 (1 ulp)
    static __device__ __forceinline__ float ex(float x) { return __expf(x); }
    static __device__ __forceinline__ float lg(float x) { return __logf(x); }
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
};

template <class T>
__host__ __device__ __forceinline__ T tp_f2_tree(T k1, T k2, T k3)
{
    // classed_version/bispectrum.cpp:629-633 (scale-free, so q stands for k)
    T c12 = T(0.5) * (k3 * k3 - k1 * k1 - k2 * k2) / (k1 * k2);
    return T(5. / 7.) + T(0.5) * c12 * (k1 / k2 + k2 / k1) + T(2. / 7.) * c12 * c12;
}

// Tabulated linear P(k), classed_version/bispectrum.cpp:787-806: linear interpolation in k between
// the bracketing entries (k_j <= k < k_j+1), the end values outside the table.
__host__ __device__ __forceinline__ double tp_pk_table(const TpPkTable &t, double k)
{
    const int n = t.nk;
    if (!(k >= t.k[0])) return t.p[0];
    if (k >= t.k[n - 1]) return t.p[n - 1];
    int c = (int)((log(k) - t.lnk0) * t.inv_dlnk);
    c = c < 0 ? 0 : (c >= t.ng ? t.ng - 1 : c);
    int j = t.idx[c];
    while (j < n - 2 && t.k[j + 1] <= k) ++j;
    const double diff = (k - t.k[j]) / (t.k[j + 1] - t.k[j]);
    return diff * t.p[j + 1] + (1 - diff) * t.p[j];
}

// Eisenstein & Hu no-wiggle linear P(k), k in h/Mpc (classed_version/bispectrum.cpp:811-841), with the
// k-independent constants folded into C; norm is C.norm except while the normalisation is computed.
// T is double in the kernels (a counting type in the operation census).
template <class T>
__host__ __device__ __forceinline__ T tp_pk_eh(const TpCosmoConst &C, T k, double norm)
{
    k = k * C.h;
    T x = 0.43 * k * C.eh_sh;
    T x2 = x * x;
    T geff = C.eh_omh2 * (C.eh_sqrt_alnu + (1.0 - C.eh_sqrt_alnu) / (1.0 + x2 * x2));
    T qeff = k / geff * C.eh_theta2;
    T L = log(2.718281828 + C.eh_cL * qeff);
    T Cq = 14.4 + 325.0 / (1.0 + 60.5 * pow(qeff, 1.11));
    T r = L / (L + Cq * qeff * qeff);
    T delk = norm * norm * pow(k * 2997.9 / C.h, 3. + C.ns) * r * r;
    T pk = 2.0 * tpd::kPi * tpd::kPi / (k * k * k) * delk;
    return C.h * C.h * C.h * pk;
}

// linear_pk (classed_version/bispectrum.cpp:785-809): the tabulated P(k) when given, else the fit
template <class T>
__host__ __device__ __forceinline__ T tp_pk_lin(const TpCosmoConst &C, T k, double norm)
{
    if constexpr (std::is_same<T, double>::value)
        if (C.tab.nk) return tp_pk_table(C.tab, k);
    return tp_pk_eh<T>(C, k, norm);
}

// Eisenstein & Hu no-wiggle P_lin at k = q / r_sigma (classed_version/bispectrum.cpp:811-841).
// The quotients of the reference are merged into two reciprocals:
//   q_eff = k theta^2 / Gamma_eff, Gamma_eff = om h^2 (sqrt(alpha_nu) + (1 - sqrt(alpha_nu)) / (1 + x^4))
//         = k theta^2 (1 + x^4) / (om h^2 (sqrt(alpha_nu) (1 + x^4) + 1 - sqrt(alpha_nu)))
//   T = L / (L + C q_eff^2), C = 14.4 + 325 / den, den = 1 + 60.5 q_eff^1.11
//     = L den / (L den + (14.4 den + 325) q_eff^2)
//   P_lin ~ Delta^2 / q^3 = exp(3 n_s (ln q + ln(2997.9 / r_sigma))) T^2 / q^3
//                        = exp(3 n_s (ln q + ln(2997.9 / r_sigma)) - 3 ln q) T^2
template <class T>
__host__ __device__ __forceinline__ T tp_plin_q(const TpNodeT<T> &N, const TpEHT<T> &E, T q, T lq)
{
    if constexpr (std::is_same<T, double>::value || std::is_same<T, float>::value)
        if (E.tab.nk) return (T)tp_pk_table(E.tab, (double)q * (double)N.inv_rs);
    T kk = q * N.inv_rs * E.h;  // 1/Mpc
    T x = E.sh043 * kk;
    T x2 = x * x;
    T d = T(1) + x2 * x2;
    T qeff = kk * E.theta2 * d * TpMath<T>::rcp(E.omh2 * (E.sqrt_alnu * d + E.one_m_sqrt_alnu));
    T L = TpMath<T>::lg(E.e_const + E.cL * qeff);
    T den = T(1) + T(60.5) * TpMath<T>::ex(T(1.11) * TpMath<T>::lg(qeff));
    T Ld = L * den;
    T r = Ld * TpMath<T>::rcp(Ld + (T(14.4) * den + T(325.0)) * qeff * qeff);
    return N.pk_pref * TpMath<T>::ex(E.three_ns * (lq + N.lnkoff) - T(3) * lq) * r * r;
}

// Scale-free per-point quantities, evaluated once per point in FP64: q_i = l_i ks_n for
// every node n, so ratios of wavenumbers, the F2 tree kernels (classed_version/
// bispectrum.cpp:629-633) and r1, r2 of eq. B8 (:179-192) do not depend on the node.
template <class T>
struct TpPointPreT {
    T l[3];    // ell1, ell2, ell3
    T ll[3];   // ln ell_i
    T f2t[3];  // F2_tree(k1,k2,k3), F2_tree(k2,k3,k1), F2_tree(k3,k1,k2)
    T lr1;     // ln r1, r1 = q_min / q_max
    T r2;      // (q_mid + q_min - q_max) / q_max
};
typedef TpPointPreT<double> TpPointPre;

template <class T>
__host__ __device__ __forceinline__ void tp_point_pre(T l1, T l2, T l3, TpPointPreT<T> *P)
{
    P->l[0] = l1;
    P->l[1] = l2;
    P->l[2] = l3;
    P->ll[0] = TpMath<T>::lg(l1);
    P->ll[1] = TpMath<T>::lg(l2);
    P->ll[2] = TpMath<T>::lg(l3);
    P->f2t[0] = tp_f2_tree<T>(l1, l2, l3);
    P->f2t[1] = tp_f2_tree<T>(l2, l3, l1);
    P->f2t[2] = tp_f2_tree<T>(l3, l1, l2);
    T a = l1, b = l2, c = l3, t;
    if (a < b) { t = a; a = b; b = t; }
    if (b < c) { t = b; b = c; c = t; }
    if (a < b) { t = a; a = b; b = t; }
    P->lr1 = TpMath<T>::lg(c / a);
    P->r2 = (b + c - a) / a;
}

// BiHalofit matter bispectrum at one z-node, classed_version/bispectrum.cpp:166-236.
// Quotients are rewritten with one reciprocal each:
//   1 / (I_1h (1 + 1/(c_n q)))  = c_n q / (I_1h (c_n q + 1))                    (eq. B4, B7)
//   (1 + f_n q^2) / (1 + g_n q + h_n q^2) D^2 P_lin + 1 / (M (1 + 1/(p_n q)^3))
//     = (A D^2 P_lin C + B (p_n q)^3) / (B C),  C = M ((p_n q)^3 + 1)            (eq. B6)
// The node value splits into the P_E term of eq. B6, a function of one wavenumber (tp_pse_node),
// and the rest, which depends on the triangle shape through r1, r2 (tp_bispec_joint); kernels that
// share wavenumbers between points evaluate the P_E terms once per distinct wavenumber.
template <class T>
__host__ __device__ __forceinline__ T tp_pse_node(const TpNodeT<T> &N, const TpEHT<T> &E, T qi, T lqi)
{
    T pl = tp_plin_q(N, E, qi, lqi);
    T pq = N.pn * qi;
    T pq3 = pq * pq * pq;
    T A = T(1) + N.fn * qi * qi;
    T B = T(1) + N.gn * qi + N.hn * qi * qi;
    T C = (N.mn * TpMath<T>::ex(N.mun * lqi) + N.nn * TpMath<T>::ex(N.nun * lqi)) * (pq3 + T(1));
    return (A * N.D1sq * pl * C + B * pq3) * TpMath<T>::rcp(B * C);  // eq. B6
}

// The parts of the one-halo term that depend on the node and the triangle shape only (eq. B7:
// a_n from r1, alpha_n and beta_n from r2); a kernel whose terms share node and shape (the Ogata sums
// of Gamma) forms them once.
template <class T>
struct TpShapeNodeT {
    T an, alphan, betan;
};

template <class T>
__host__ __device__ __forceinline__ void tp_shape_node(const TpNodeT<T> &N, T lr1, T r2, TpShapeNodeT<T> *S)
{
    S->an = TpMath<T>::ex(N.lnA0 + N.m031ln10 * TpMath<T>::ex(N.c1 * lr1));   // eq. B7
    T alphan = TpMath<T>::ex(N.a0ln + N.a1ln * r2 * r2);
    if (alphan > N.acap) alphan = N.acap;
    S->alphan = alphan;
    S->betan = TpMath<T>::ex(N.b0ln + N.b1ln * r2);
}

template <class T>
__host__ __device__ __forceinline__ T tp_bispec_joint_s(const TpNodeT<T> &N, const TpShapeNodeT<T> &S, const T q[3],
                                                        const T lq[3], const T f2t[3], const T pse[3])
{
    T bs1h = T(1), damp = T(1);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const T qi = q[i], lqi = lq[i];
        T oneh = S.an * TpMath<T>::ex(S.alphan * lqi) + N.bn * TpMath<T>::ex(S.betan * lqi);   // eq. B4
        T cq = N.cn * qi;
        bs1h *= cq * TpMath<T>::rcp(oneh * (cq + T(1)));
        damp *= T(1) + N.en * qi;
    }
    T bs3h = T(2) * ((f2t[0] + N.dn * q[2]) * pse[0] * pse[1] + (f2t[1] + N.dn * q[0]) * pse[1] * pse[2] +
                     (f2t[2] + N.dn * q[1]) * pse[2] * pse[0]);                            // eq. B5 with F2 (:614-627)
    return bs1h + bs3h * TpMath<T>::rcp(damp);
}

template <class T>
__host__ __device__ __forceinline__ T tp_bispec_joint(const TpNodeT<T> &N, const T q[3], const T lq[3], const T f2t[3],
                                                      T lr1, T r2, const T pse[3])
{
    TpShapeNodeT<T> S;
    tp_shape_node<T>(N, lr1, r2, &S);
    return tp_bispec_joint_s<T>(N, S, q, lq, f2t, pse);
}

template <class T>
__host__ __device__ __forceinline__ T tp_bispec_node(const TpNodeT<T> &N, const TpEHT<T> &E, const T q[3], const T lq[3],
                                            const T f2t[3], T lr1, T r2)
{
    T pse[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) pse[i] = tp_pse_node<T>(N, E, q[i], lq[i]);
    return tp_bispec_joint<T>(N, q, lq, f2t, lr1, r2, pse);
}

// Same bispectrum from raw wavenumbers (used where the per-point hoisting does not apply)
template <class T>
__device__ __forceinline__ T tp_bispec_q(const TpNodeT<T> &N, const TpEHT<T> &E, T q1, T q2, T q3)
{
    T a = q1, b = q2, c = q3, t;
    if (a < b) { t = a; a = b; b = t; }
    if (b < c) { t = b; b = c; c = t; }
    if (a < b) { t = a; a = b; b = t; }
    T q[3] = {q1, q2, q3};
    T lq[3] = {TpMath<T>::lg(q1), TpMath<T>::lg(q2), TpMath<T>::lg(q3)};
    T f2t[3] = {tp_f2_tree(q1, q2, q3), tp_f2_tree(q2, q3, q1), tp_f2_tree(q3, q1, q2)};
    return tp_bispec_node<T>(N, E, q, lq, f2t, TpMath<T>::lg(c / a), (b + c - a) / a);
}

// Limber sum over NN z-nodes (FP64 accumulation), without the zero-ell guard
template <class T, int NN>
__device__ __forceinline__ double tp_bkappa_sum_n(const TpNodeT<T> *__restrict__ nodes, const TpEHT<T> &E,
                                                  double l1, double l2, double l3)
{
    TpPointPre P;
    tp_point_pre<double>(l1, l2, l3, &P);
    const T L[3] = {(T)l1, (T)l2, (T)l3};
    const T LL[3] = {(T)P.ll[0], (T)P.ll[1], (T)P.ll[2]};
    const T F2T[3] = {(T)P.f2t[0], (T)P.f2t[1], (T)P.f2t[2]};
    const T LR1 = (T)P.lr1, R2 = (T)P.r2;
    double acc = 0.0;
#pragma unroll 2
    for (int n = 0; n < NN; ++n) {
        const TpNodeT<T> &N = nodes[n];
        const T ks = N.ks, lks = N.lnks;
        T q[3] = {L[0] * ks, L[1] * ks, L[2] * ks};
        T lq[3] = {LL[0] + lks, LL[1] + lks, LL[2] + lks};
        T b = tp_bispec_node<T>(N, E, q, lq, F2T, LR1, R2);
        acc += (double)(N.wz * b);
    }
    return acc;
}

// One node of the Ogata sums of classed_version/gamma.cpp:298-314 with the per-point shape
// hoisted: every term of a cubature point (z, phi, psi) has the sides
//   l = (psi_k / A) (cos psi, sin psi, sqrt(1 + sin(2 psi) cos phi)),
// i.e. one triangle shape P (unit scale) at the scale s = psi_k / A.  ln_s = ln psi_k - ln A.
// S: tp_shape_node(N, P.lr1, P.r2), the same for every term of a point.
template <class T>
__host__ __device__ __forceinline__ T tp_ogata_term(const TpNodeT<T> &N, const TpEHT<T> &E, const TpPointPreT<T> &P,
                                                     const TpShapeNodeT<T> &S, T s, T ln_s, T prodk)
{
    const T sk = s * N.ks, lsk = ln_s + N.lnks;
    T q[3] = {P.l[0] * sk, P.l[1] * sk, P.l[2] * sk};
    T lq[3] = {P.ll[0] + lsk, P.ll[1] + lsk, P.ll[2] + lsk};
    T pse[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) pse[i] = tp_pse_node<T>(N, E, q[i], lq[i]);
    return N.wz * tp_bispec_joint_s<T>(N, S, q, lq, P.f2t, pse) * prodk;
}

// The 96-node Limber sum of the reference's bkappa (classed_version/bispectrum.cpp:238-286),
// accumulated in a fixed order: 8 partial sums of 12 consecutive nodes, combined pairwise,
// ((p0 + p1) + (p2 + p3)) + ((p4 + p5) + (p6 + p7)).  A point whose nodes are split over
// 1, 2, 4 or 8 waves forms the same tree, so every launch shape gives the same bits.
constexpr int TP_NPART = 8, TP_PART_NODES = TP_NNODES / TP_NPART;

// Adds partial k (0-based, in order) to the pairwise tree: s1, s2, s4 hold the pending left operands
// of the levels (p0 + p1), ((p0 + p1) + (p2 + p3)), ...; returns the running subtree value, after
// partial 7 the whole tree ((p0 + p1) + (p2 + p3)) + ((p4 + p5) + (p6 + p7)).
__device__ __forceinline__ double tp_tree8_push(int k, double p, double &s1, double &s2, double &s4)
{
    double v = p;
    if (k & 1) {
        v = s1 + v;
        if (k & 2) {
            v = s2 + v;
            if (k & 4)
                v = s4 + v;
            else
                s4 = v;
        } else {
            s2 = v;
        }
    } else {
        s1 = v;
    }
    return v;
}

// partials [part0, part0 + NP) of the tree, combined pairwise (NP a power of two)
template <class T, int NP>
__device__ __forceinline__ double tp_bkappa_parts(const TpNodeT<T> *__restrict__ nodes, const TpEHT<T> &E,
                                                  const TpPointPre &P, int part0)
{
    const T L[3] = {(T)P.l[0], (T)P.l[1], (T)P.l[2]};
    const T LL[3] = {(T)P.ll[0], (T)P.ll[1], (T)P.ll[2]};
    const T F2T[3] = {(T)P.f2t[0], (T)P.f2t[1], (T)P.f2t[2]};
    const T LR1 = (T)P.lr1, R2 = (T)P.r2;
    // one pass over the nodes; at the end of each partial its sum joins the pairwise tree
    double s1 = 0.0, s2 = 0.0, s4 = 0.0, v = 0.0, acc = 0.0;
    const int n0 = part0 * TP_PART_NODES, n1 = n0 + NP * TP_PART_NODES;
#pragma unroll 2
    for (int n = n0; n < n1; ++n) {
        const TpNodeT<T> &N = nodes[n];
        const T ks = N.ks, lks = N.lnks;
        T q[3] = {L[0] * ks, L[1] * ks, L[2] * ks};
        T lq[3] = {LL[0] + lks, LL[1] + lks, LL[2] + lks};
        T b = tp_bispec_node<T>(N, E, q, lq, F2T, LR1, R2);
        acc += (double)(N.wz * b);
        const int m = n - n0 + 1;
        if (m % TP_PART_NODES == 0) {  // partial k = m / 12 - 1 complete
            v = tp_tree8_push(m / TP_PART_NODES - 1, acc, s1, s2, s4);
            acc = 0.0;
        }
    }
    return v;
}

template <class T>
__device__ __forceinline__ double tp_bkappa_sum(const TpNodeT<T> *__restrict__ nodes, const TpEHT<T> &E,
                                                double l1, double l2, double l3)
{
    TpPointPre P;
    tp_point_pre<double>(l1, l2, l3, &P);
    return tp_bkappa_parts<T, TP_NPART>(nodes, E, P, 0);
}

// the reference's test_analytical B_kappa (classed_version/bispectrum.cpp:242-255, 989-996)
__device__ __forceinline__ double tp_bkappa_analytic(double l1, double l2, double l3)
{
    const double wts[9] = {0, 1.0e+8, 1.0e+6, 1.0e+4, 1.0e+2, 1.0e+0, 5.0e-3, 1.0e-5, 1.0e-6};
    const double avs[9] = {3.0e+3, 1.0e+4, 3.0e+4, 1.0e+5, 3.0e+5, 1.0e+6, 3.0e+6, 1.0e+7, 3.0e+7};
    // cos(acos(x)) with x clamped to [-1, 1]: rounding can push |x| past 1 at phi = 0, pi
    double cphi = (l1 * l1 + l2 * l2 - l3 * l3) / (2 * l1 * l2);
    cphi = fmin(1.0, fmax(-1.0, cphi));
    double s = 0;
    for (int i = 1; i < 9; ++i) {
        double a = avs[i];
        double a2 = a * a, a4 = a2 * a2;
        double bis_pref = -(tpd::kPi * tpd::kPi) / (8 * 72 * (a4 * a4));
        double pref = (l1 * l1) * (l2 * l2) * (l1 * l1 + l2 * l2 + 2 * l1 * l2 * cphi);
        double ex = exp(-(l1 * l1 + l2 * l2 + l1 * l2 * cphi) / (6 * a));
        s += wts[i] * (bis_pref * pref * ex);
    }
    return s * (3. * (2 * tpd::kPi) * (2 * tpd::kPi));
}

template <class T>
__host__ __device__ __forceinline__ T tp_uhat(T eta)
{
    T t = T(0.5) * eta * eta;  // classed_version/apertureStatistics.cpp:10-14
    return t * TpMath<T>::ex(-t);
}

// l3 exactly as the reference forms it (no contraction: l1 = l2, phi = pi gives 0)
__device__ __forceinline__ double tp_ell3(double l1, double l2, double phi)
{
#pragma clang fp contract(off)
    double s = l1 * l1 + l2 * l2 + 2 * l1 * l2 * cos(phi);
    return sqrt(s);
}

// The <Map^3> integrand (classed_version/apertureStatistics.cpp:16-28) in two halves around the
// z-sum, for kernels that split a point's nodes over several waves.  tp_map3_begin returns
// TP_PT_DONE for a non-finite ell (flagged; the value is NaN), TP_PT_NOSUM when B_kappa needs no
// z-sum (an ell is 0, or the analytic test bispectrum) and TP_PT_SUM otherwise (P filled).
enum { TP_PT_DONE = 0, TP_PT_NOSUM = 1, TP_PT_SUM = 2 };
__device__ __forceinline__ int tp_map3_begin(double l1, double l2, double phi, int mode, int *flag, TpPointPre *P)
{
    double l3 = tp_ell3(l1, l2, phi);
    P->l[0] = l1;
    P->l[1] = l2;
    P->l[2] = l3;
    if (!isfinite(l1) || !isfinite(l2) || !isfinite(l3)) {
        atomicOr(flag, 1);
        return TP_PT_DONE;
    }
    if (l1 == 0 || l2 == 0 || l3 == 0 || mode == TP_MODE_ANALYTIC) return TP_PT_NOSUM;  // bispectrum.cpp:240-241
    tp_point_pre<double>(l1, l2, l3, P);
    return TP_PT_SUM;
}
__device__ __forceinline__ double tp_map3_end(const TpPointPre &P, int kind, double bk_sum, double th1, double th2,
                                              double th3, int *flag)
{
    if (kind == TP_PT_DONE) return __builtin_nan("");
    const double l1 = P.l[0], l2 = P.l[1], l3 = P.l[2];
    double bk;
    if (kind == TP_PT_SUM)
        bk = bk_sum;
    else if (l1 == 0 || l2 == 0 || l3 == 0)
        bk = 0.0;
    else
        bk = tp_bkappa_analytic(l1, l2, l3);  // TP_MODE_ANALYTIC
    if (!isfinite(bk)) atomicOr(flag, 2);
    return l1 * l2 * bk * tp_uhat<double>(l1 * th1) * tp_uhat<double>(l2 * th2) * tp_uhat<double>(l3 * th3);
}

template <class T>
__device__ __forceinline__ double tp_map3_integrand(const TpNodeT<T> *__restrict__ nodes, const TpEHT<T> &E,
                                                    double l1, double l2, double phi, double th1, double th2,
                                                    double th3, int mode, int *flag)
{
    TpPointPre P;
    const int kind = tp_map3_begin(l1, l2, phi, mode, flag, &P);
    const double s = kind == TP_PT_SUM ? tp_bkappa_parts<T, TP_NPART>(nodes, E, P, 0) : 0.0;
    return tp_map3_end(P, kind, s, th1, th2, th3, flag);
}

// --------------------------------------------------- Genz-Malik 3-D rule
// Point p of a region, in the reference's order (classed_version/hcubature.cpp:
// 343-414, 476-503): center; -+lambda2, -+lambda4 on each axis; (lambda4,lambda4)
// pairs; 2^3 lambda5 corners in Gray-code order.  code: 0 centre, +-1 lambda2,
// +-2 lambda4, +-3 lambda5.
__constant__ signed char c_gm3_code[TP_NPTS_GM3][3] = {
    {0, 0, 0},
    {-1, 0, 0}, {1, 0, 0}, {-2, 0, 0}, {2, 0, 0},
    {0, -1, 0}, {0, 1, 0}, {0, -2, 0}, {0, 2, 0},
    {0, 0, -1}, {0, 0, 1}, {0, 0, -2}, {0, 0, 2},
    {-2, -2, 0}, {2, -2, 0}, {2, 2, 0}, {-2, 2, 0},
    {-2, 0, -2}, {2, 0, -2}, {2, 0, 2}, {-2, 0, 2},
    {0, -2, -2}, {0, 2, -2}, {0, 2, 2}, {0, -2, 2},
    {3, 3, 3}, {-3, 3, 3}, {-3, -3, 3}, {3, -3, 3},
    {3, -3, -3}, {-3, -3, -3}, {-3, 3, -3}, {3, 3, -3}};

__device__ __forceinline__ double tp_gm_coord(double c, double hw, int code)
{
#pragma clang fp contract(off)
    const double l2 = 0.3585685828003180919906451539079374954541;  // sqrt(9/70)
    const double l4 = 0.9486832980505137995996680633298155601160;  // sqrt(9/10)
    const double l5 = 0.6882472016116852977216287342936235251269;  // sqrt(9/19)
    int a = code < 0 ? -code : code;
    if (a == 0) return c;
    double lam = (a == 1) ? l2 : (a == 2 ? l4 : l5);
    double w = hw * lam;
    return code < 0 ? c - w : c + w;
}

__device__ __forceinline__ void gm_point3(const TpRegion &R, int p, double *x0, double *x1, double *x2)
{
    *x0 = tp_gm_coord(R.c[0], R.hw[0], c_gm3_code[p][0]);
    *x1 = tp_gm_coord(R.c[1], R.hw[1], c_gm3_code[p][1]);
    *x2 = tp_gm_coord(R.c[2], R.hw[2], c_gm3_code[p][2]);
}

// value, error and split dimension of one 3-D region from its 33 values
// (classed_version/hcubature.cpp:515-575, fdim = 1)
__device__ __forceinline__ void gm_rule3(double vol, const double *v, TpRegionOut *out)
{
#pragma clang fp contract(off)
    const double l2 = 0.3585685828003180919906451539079374954541;
    const double l4 = 0.9486832980505137995996680633298155601160;
    const double ratio = (l2 * l2) / (l4 * l4);
    const double w1 = -10936.0 / 19683.0, w3 = 620.0 / 19683.0, w5 = 6859.0 / 19683.0 / 8.0;
    const double wE1 = -1671.0 / 729.0, wE3 = -35.0 / 1458.0;
    const double wt2 = 980. / 6561., wt4 = 200. / 19683., wtE2 = 245. / 486., wtE4 = 25. / 729.;
    double val0 = v[0], sum2 = 0, sum3 = 0, sum4 = 0, sum5 = 0;
    double diff[3];
    for (int k = 0; k < 3; ++k) {
        double a0 = v[1 + 4 * k], a1 = v[2 + 4 * k], a2 = v[3 + 4 * k], a3 = v[4 + 4 * k];
        sum2 += a0 + a1;
        sum3 += a2 + a3;
        diff[k] = fabs(a0 + a1 - 2 * val0 - ratio * (a2 + a3 - 2 * val0));
    }
    for (int k = 0; k < 12; ++k) sum4 += v[13 + k];
    for (int k = 0; k < 8; ++k) sum5 += v[25 + k];
    double result = vol * (w1 * val0 + wt2 * sum2 + w3 * sum3 + wt4 * sum4 + w5 * sum5);
    double res5th = vol * (wE1 * val0 + wtE2 * sum2 + wE3 * sum3 + wtE4 * sum4);
    out->val = result;
    out->err = fabs(res5th - result);
    out->corr = 0.0;
    out->cvar = 0.0;
    int dd = 0;
    double md = 0;
    for (int i = 0; i < 3; ++i)
        if (diff[i] > md) {
            md = diff[i];
            dd = i;
        }
    out->split = dd;
    out->flag = (isfinite(result) && isfinite(res5th)) ? 0 : 1;
}
