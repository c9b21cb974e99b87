// Device functions of the InteraxFF pair energies: Tang-Toennies damped C6/C8/C10 dispersion, Slater-overlap
// exchange and the short-range correction, with their radial derivatives.  Templated over the arithmetic type:
// float for the single and mixed precision modes, double for the double precision mode.
//
// The mathematics is that of openmmapi/include/internal/InteraxPairTerms.h:167-337 (docs/SPEC.md:37-122); the
// evaluation is organised for the VALU of gfx950: reciprocals, square roots and exponentials use the hardware
// approximations (v_rcp_f32, v_rsq_f32, v_exp_f32) in FP32, both exponentials of a Slater overlap are shared
// between its closed form and its Taylor branch, polynomials are evaluated in Horner form with compile-time
// coefficients, and everything that depends only on a pair of particle types can come from a table (PairTables).
#ifndef IXH_PAIR_MATH_H
#define IXH_PAIR_MATH_H

#include "ixh_common.h"

namespace ixh {

constexpr double LOG2E = 1.4426950408889634;
constexpr double INV_PI = 0.31830988618379067154;
constexpr double INV_8PI = 0.039788735772973833942;
constexpr double EIGHT_PI = 25.132741228718345908;

// Taylor expansion of the overlap: pi l^3 exp(x) S(x l; l, l(1+t)) = sum_n t^n sum_k V[n][k] x^k, exact rationals
// (the same table as InteraxPairTerms.h:96-104; it is a property of the integral).  Row n has terms k <= n+2.
struct OverlapTaylor {
    static constexpr int ORDER = 6;
    static constexpr int DEGREE = ORDER+2;
    __host__ __device__ static constexpr double v(int n, int k) {
        constexpr double table[7][9] = {
            {1.0/64, 1.0/64, 1.0/192, 0, 0, 0, 0, 0, 0},
            {-3.0/128, -3.0/128, -1.0/192, 1.0/384, 0, 0, 0, 0, 0},
            {3.0/128, 3.0/128, 1.0/768, -5.0/768, 1.0/1280, 0, 0, 0, 0},
            {-5.0/256, -5.0/256, 1.0/256, 1.0/96, -11.0/3840, 1.0/5760, 0, 0, 0},
            {15.0/1024, 15.0/1024, -9.0/1024, -7.0/512, 33.0/5120, -19.0/23040, 1.0/32256, 0, 0},
            {-21.0/2048, -21.0/2048, 13.0/1024, 33.0/2048, -59.0/5120, 107.0/46080, -29.0/161280, 1.0/215040, 0},
            {7.0/1024, 7.0/1024, -191.0/12288, -73.0/4096, 369.0/20480, -233.0/46080, 389.0/645120, -41.0/1290240, 1.0/1658880}};
        return table[n][k];
    }
    // c_k(t) = sum_{n <= order} V[n][k] t^n: the coefficient of x^k once t is fixed
    __host__ static double coefficient(int order, int k, double t) {
        double c = 0;
        for (int n = order; n >= 0; n--)
            c = c*t + v(n, k);
        return c;
    }
};

// Where the Taylor branch is used (InteraxPairTerms.h:132-160).
struct OverlapScheme {
    int order;              // 6 (Exact) or 2 (InteraxOrder2)
    int about_smaller;      // 1: expand about the smaller width; 0: about the width of the lower-index particle
    float max_delta_f, max_t_f, max_tx_f;
    double max_delta, max_t, max_tx;
};

template <typename Real>
struct SchemeValues;
template <>
struct SchemeValues<float> {
    __device__ static float max_delta(const OverlapScheme& s) { return s.max_delta_f; }
    __device__ static float max_t(const OverlapScheme& s) { return s.max_t_f; }
    __device__ static float max_tx(const OverlapScheme& s) { return s.max_tx_f; }
};
template <>
struct SchemeValues<double> {
    __device__ static double max_delta(const OverlapScheme& s) { return s.max_delta; }
    __device__ static double max_t(const OverlapScheme& s) { return s.max_t; }
    __device__ static double max_tx(const OverlapScheme& s) { return s.max_tx; }
};

// P(x) = sum_k c_k(t) x^k of the Taylor branch and its x derivative, for expansion order ORDER.  The coefficients
// c_k(t) are formed first (Horner in t, with the t^(k-2) factor of the rows k >= 3 pulled out), then P and P' by a
// joint Horner recurrence in x: about 60 FMAs at order 6 instead of the 98 of a row-by-row expansion.
template <typename Real, int ORDER>
__device__ __forceinline__ void taylor_polynomial(Real t, Real x, Real& P, Real& dP) {
    constexpr int K = ORDER+2;
    Real tp[ORDER+1];
    tp[0] = 1;
#pragma unroll
    for (int n = 1; n <= ORDER; n++)
        tp[n] = tp[n-1]*t;
    Real c[K+1];
#pragma unroll
    for (int k = 0; k <= K; k++) {
        const int n0 = (k > 2 ? k-2 : 0);
        Real acc = Real(OverlapTaylor::v(ORDER, k));
#pragma unroll
        for (int n = ORDER-1; n >= n0; n--)
            acc = acc*t + Real(OverlapTaylor::v(n, k));
        c[k] = (n0 > 0 ? acc*tp[n0] : acc);
    }
    P = c[K];
    dP = 0;
#pragma unroll
    for (int k = K-1; k >= 0; k--) {
        dP = dP*x + P;
        P = P*x + c[k];
    }
}

// Slater overlap S(r; a, b) and dS/dr.  ia, ib are 1/a, 1/b.  For the InteraxOrder2 scheme a must be the width of
// the particle with the lower index.
template <typename Real>
__device__ __forceinline__ void slater_overlap(Real r, Real invR, Real a, Real ia, Real b, Real ib, const OverlapScheme& scheme,
                                               Real& overlap, Real& deriv) {
    const Real log2e = Real(LOG2E);
    Real expA = fast_exp2(-r*ia*log2e);
    Real expB = fast_exp2(-r*ib*log2e);

    // Taylor expansion about l (the smaller width, or a).
    bool aSmaller = (a <= b);
    bool aboutA = (scheme.about_smaller ? aSmaller : true);
    Real invL = (aboutA ? ia : ib);
    Real expL = (aboutA ? expA : expB);
    Real delta = (aboutA ? b-a : a-b);
    Real t = delta*invL;
    Real x = r*invL;
    Real absT = fabs(t);
    bool useTaylor = (fabs(delta) < SchemeValues<Real>::max_delta(scheme)) ||
                     ((absT < SchemeValues<Real>::max_t(scheme)) && (absT*x < SchemeValues<Real>::max_tx(scheme)));

    Real taylorS = 0, taylorD = 0;
    if (__builtin_expect(useTaylor, 1)) {
        Real P, dP;
        if (scheme.order == 2)
            taylor_polynomial<Real, 2>(t, x, P, dP);
        else
            taylor_polynomial<Real, 6>(t, x, P, dP);
        Real prefactor = expL*invL*invL*invL*Real(INV_PI);
        taylorS = prefactor*P;
        taylorD = prefactor*invL*(dP-P);
    }
    Real closedS = 0, closedD = 0;
    if (!useTaylor) {
        // S = F/(8 pi r), F = 4 a^2 b^2/D^3 (e_a - e_b) + r/D^2 (a e_a + b e_b), D = b^2 - a^2
        Real aa = a*a, bb = b*b;
        Real d = bb-aa;
        Real invD2 = fast_rcp(d*d);
        Real c = Real(4)*aa*bb*invD2*(d*invD2);
        Real sum = a*expA+b*expB;
        Real f = c*(expA-expB) + r*invD2*sum;
        Real dfdr = c*(expB*ib-expA*ia) + invD2*(sum-r*(expA+expB));
        closedS = f*invR*Real(INV_8PI);
        closedD = (r*dfdr-f)*invR*invR*Real(INV_8PI);
    }
    overlap = (useTaylor ? taylorS : closedS);
    deriv = (useTaylor ? taylorD : closedD);
}

// Number of terms of the Tang-Toennies series tail sum_{k>10} x^k/k! evaluated below x = 5.5: the reference sums to
// k = 40 (InteraxPairTerms.h:240-244); in FP32 the terms beyond k = 26 change the tail by less than 2e-9 relative.
template <typename Real>
struct TangToenniesTail {
    static constexpr int TERMS = 30;
};
template <>
struct TangToenniesTail<float> {
    static constexpr int TERMS = 16;
};

// 10!/(10+m)!, the coefficient of x^m in tail/(x^10/10!)
__host__ __device__ constexpr double tail_coefficient(int m) {
    double c = 1;
    for (int j = 1; j <= m; j++)
        c /= (10+j);
    return c;
}

template <typename Real, int M, int K>
struct TailPolynomial {   // sum_{m=K}^{M} c_m x^(m-K)
    static constexpr double C = tail_coefficient(K);
    __device__ static __forceinline__ Real eval(Real x) { return Real(C) + x*TailPolynomial<Real, M, K+1>::eval(x); }
};
template <typename Real, int M>
struct TailPolynomial<Real, M, M> {
    static constexpr double C = tail_coefficient(M);
    __device__ static __forceinline__ Real eval(Real) { return Real(C); }
};

// Tang-Toennies f_6, f_8, f_10 and df/dx (InteraxPairTerms.h:223-253).  Below x = (n+1)/2 the closed form
// 1 - exp(-x) sum_{k<=n} x^k/k! cancels and f_n = exp(-x) sum_{k>n} x^k/k! is used, a sum of positive terms.
// In two steps: damping_main (branch-free: the closed forms and the derivatives) and damping_tail (the series below
// x = 5.5, the only divergent step), so that a lane's several pairs can share one block of straight-line code.
template <typename Real>
struct Damping {
    Real x, expX, t6, t7, t8, t9, t10;   // t_n = x^n/n!
    Real f[3], df[3];                    // df_n = exp(-x) t_n
};

template <typename Real>
__device__ __forceinline__ void damping_main(Real x, Damping<Real>& d) {
    const Real expX = fast_exp2(-x*Real(LOG2E));
    // s6 = sum_{k<=6} x^k/k! in Horner form; t_n = x^n/n!
    Real s6 = Real(1.0/720);
    s6 = s6*x + Real(1.0/120);
    s6 = s6*x + Real(1.0/24);
    s6 = s6*x + Real(1.0/6);
    s6 = s6*x + Real(0.5);
    s6 = s6*x + Real(1);
    s6 = s6*x + Real(1);
    const Real x2 = x*x;
    const Real x3 = x2*x;
    const Real t6 = x3*x3*Real(1.0/720);
    const Real t7 = t6*x*Real(1.0/7);
    const Real t8 = t7*x*Real(1.0/8);
    const Real t9 = t8*x*Real(1.0/9);
    const Real t10 = t9*x*Real(1.0/10);
    const Real s8 = s6+(t7+t8);
    const Real s10 = s8+(t9+t10);
    d.x = x;
    d.expX = expX;
    d.t6 = t6;
    d.t7 = t7;
    d.t8 = t8;
    d.t9 = t9;
    d.t10 = t10;
    d.f[0] = 1-expX*s6;
    d.f[1] = 1-expX*s8;
    d.f[2] = 1-expX*s10;
    d.df[0] = expX*t6;
    d.df[1] = expX*t8;
    d.df[2] = expX*t10;
}

template <typename Real>
__device__ __forceinline__ void damping_tail(Damping<Real>& d) {
    const Real x = d.x;
    if (x < Real(5.5)) {
        const Real tail10 = d.t10*x*TailPolynomial<Real, TangToenniesTail<Real>::TERMS, 1>::eval(x);
        const Real tail8 = tail10+(d.t9+d.t10);
        const Real tail6 = tail8+(d.t7+d.t8);
        d.f[2] = d.expX*tail10;
        if (x < Real(4.5))
            d.f[1] = d.expX*tail8;
        if (x < Real(3.5))
            d.f[0] = d.expX*tail6;
    }
}

template <typename Real>
__device__ __forceinline__ void tang_toennies(Real x, Real f[3], Real df[3]) {
    Damping<Real> d;
    damping_main(x, d);
    damping_tail(d);
    for (int k = 0; k < 3; k++) {
        f[k] = d.f[k];
        df[k] = d.df[k];
    }
}

// Per-particle parameters of a fused pair force, in the arithmetic type.
template <typename Real>
struct PairParticle {
    Real m1, m2, m3, alpha;     // dispersion
    Real wd;                    // dispersion width
    Real amp, wx, iwx;          // exchange amplitude, width, 1/width
    Real k, l, ws, iws;         // short range K, L, width, 1/width
};

// Settings shared by every pair of a launch.
struct PairSettings {
    OverlapScheme exchange_scheme, short_range_scheme;
    float damping_scale_f;
    double damping_scale;
    int shared_overlap;   // exchange and short range use identical widths and schemes: one overlap per pair
};

// Energy and dE/dr of the selected terms of a pair at distance r.  ordered: whether particle i has the lower
// original index (InteraxOrder2 expands about the width of that particle).
template <typename Real, int TERMS>
__device__ __forceinline__ void pair_energy(Real r, Real invR, const PairParticle<Real>& pi, const PairParticle<Real>& pj,
                                            bool iFirst, const PairSettings& set, Real& energy, Real& dEdR) {
    energy = 0;
    dEdR = 0;
    if (TERMS & IXH_TERM_DISPERSION) {
        // InteraxPairTerms.h:267-293, gate ReferenceInteraxKernels.cpp:87-88
        if (pi.alpha*pj.alpha != Real(0)) {
            Real den = pi.alpha*pj.m1 + pj.alpha*pi.m1;
            den = (den > Real(1e-30) ? den : Real(1e-30));
            Real pre = pi.alpha*pj.alpha*fast_rcp(den);
            Real c6 = pre*pi.m1*pj.m1;
            Real c8 = Real(1.5)*pre*(pi.m1*pj.m2 + pi.m2*pj.m1);
            Real c10 = pre*(Real(2)*(pi.m1*pj.m3 + pi.m3*pj.m1) + Real(4.2)*pi.m2*pj.m2);
            Real beta = (sizeof(Real) == 4 ? Real(set.damping_scale_f) : Real(set.damping_scale));
            Real dxdr = Real(2)*beta*fast_rcp(pi.wd+pj.wd);
            Real f[3], df[3];
            tang_toennies<Real>(dxdr*r, f, df);
            Real invR2 = invR*invR;
            Real invR6 = invR2*invR2*invR2;
            Real e6 = c6*invR6;
            Real e8 = c8*invR6*invR2;
            Real e10 = c10*invR6*invR2*invR2;
            energy -= f[0]*e6 + f[1]*e8 + f[2]*e10;
            dEdR -= e6*(df[0]*dxdr - Real(6)*f[0]*invR) + e8*(df[1]*dxdr - Real(8)*f[1]*invR) +
                    e10*(df[2]*dxdr - Real(10)*f[2]*invR);
        }
    }
    constexpr bool EXCH = (TERMS & IXH_TERM_EXCHANGE) != 0;
    constexpr bool SR = (TERMS & IXH_TERM_SHORT_RANGE) != 0;
    if (EXCH || SR) {
        // InteraxPairTerms.h:305-337 with the gates and index order of ReferenceInteraxKernels.cpp:130-136, 173-179
        Real sx = 0, dsx = 0;
        bool haveX = false;
        if (EXCH && pi.wx*pj.wx != Real(0)) {
            if (iFirst || set.exchange_scheme.about_smaller)
                slater_overlap<Real>(r, invR, pi.wx, pi.iwx, pj.wx, pj.iwx, set.exchange_scheme, sx, dsx);
            else
                slater_overlap<Real>(r, invR, pj.wx, pj.iwx, pi.wx, pi.iwx, set.exchange_scheme, sx, dsx);
            Real amp = pi.amp*pj.amp;
            energy += amp*sx;
            dEdR += amp*dsx;
            haveX = true;
        }
        if (SR && pi.ws*pj.ws != Real(0)) {
            Real ss, dss;
            if (EXCH && set.shared_overlap && haveX) {
                ss = sx;
                dss = dsx;
            }
            else if (iFirst || set.short_range_scheme.about_smaller)
                slater_overlap<Real>(r, invR, pi.ws, pi.iws, pj.ws, pj.iws, set.short_range_scheme, ss, dss);
            else
                slater_overlap<Real>(r, invR, pj.ws, pj.iws, pi.ws, pi.iws, set.short_range_scheme, ss, dss);
            Real wsum = pi.ws+pj.ws;
            Real norm = Real(EIGHT_PI)*wsum*wsum*wsum;
            Real st = norm*ss;
            Real kk = Real(0.5)*(pi.k+pj.k);
            Real ll = Real(0.5)*(pi.l+pj.l);
            energy += (kk+ll*st)*st;
            dEdR += (kk+Real(2)*ll*st)*norm*dss;
        }
    }
}

// ------------------------------------------------------------------------------------------------------------
// Type tables.  Force fields assign parameters per atom type, so the pair quantities that depend only on the two
// types -- the combined C6/C8/C10 and every r-independent constant of a Slater overlap -- are computed once on
// the host in FP64 and read from LDS.  Used in FP32 when the widths and dispersion parameters fall into few
// classes; the per-pair path above handles everything else.
// ------------------------------------------------------------------------------------------------------------

constexpr int MAX_DISPERSION_CLASSES = 16;
constexpr int MAX_WIDTH_CLASSES = 8;

// One representation of S(r; a, b) for an ordered pair of classes (Slater width, amplitude or K and L), in a form
// shared by both branches:
//   S(r) = e1(r) (m1/r + sum_{k<=8} p_k r^k) + e2(r) (q1/r + q0),   e_i(r) = exp(-r/w_i).
// The Taylor branch about l (InteraxPairTerms.h:189-208) is e1 = exp(-r/l), p_k = c_k(t)/(pi l^3 l^k) and no e2 part;
// the closed form (:176-187) is S = [4a^2b^2/D^3 (e_a - e_b) + r/D^2 (a e_a + b e_b)]/(8 pi r), i.e.
// m1 = 4a^2b^2/(8 pi D^3), p_0 = a/(8 pi D^2), q1 = -m1, q0 = b/(8 pi D^2).  Every pair then runs the same
// instructions whichever branch its widths select: no divergence between Taylor and closed-form pairs of a wave.
// The fields are laid out for 16-byte LDS reads: the first 48 bytes hold everything a row of degree 2 with one
// overlap per pair needs (a1, a2: the energy S (a1 + a2 S) of its exchange and short-range terms together).
struct alignas(16) OverlapRow {
    float k1, k2;            // log2(e)/w1, log2(e)/w2
    float n1, n2;            // 1/w1, 1/w2
    float m1;
    float p[3];              // p_0 .. p_2
    float q1, q0;
    float a1, a2;            // A_i A_j + (K_i+K_j)/2 norm, (L_i+L_j)/2 norm^2 (FP64, rounded once)
    float p3;
    float amplitude;         // exchange: A_i A_j
    float norm;              // 8 pi (a+b)^3, the short-range normalisation
    float kk, ll;            // short range: (K_i+K_j)/2, (L_i+L_j)/2
    float kTaylor;           // the Taylor row is used where r*kTaylor < txLimit, the closed-form row elsewhere
    float txLimit;
    float high;              // 1 if any of p_4 .. p_8 is non-zero (a Taylor row of unequal widths), else 0
    float ph[5];             // p_4 .. p_8
    float gate;              // 0 if a or b is zero: the pair has no such term
    float pad[2];
    float coefficient(int k) const { return k < 3 ? p[k] : k == 3 ? p3 : ph[k-4]; }
    void set_coefficient(int k, float v) {
        if (k < 3)
            p[k] = v;
        else if (k == 3)
            p3 = v;
        else
            ph[k-4] = v;
    }
};
static_assert(sizeof(OverlapRow) == 28*sizeof(float), "OverlapRow layout");
// Per ordered class pair: the Taylor row, then the closed-form row (both identical when one branch is never taken),
// then 4 floats of padding: class pairs 60 words apart start in different LDS banks, so the lanes of a wave that
// read different class pairs' rows (ds_read_b128) do not conflict.
constexpr int OVERLAP_ROW_FLOATS = 28;
constexpr int OVERLAP_PAIR_FLOATS = 2*OVERLAP_ROW_FLOATS+4;

// Per-particle values the tables do not cover (FP32): the dispersion width (per atom) and the class indices.
struct TableParticle {
    float wd;
    int cls;                 // dispersion class | exchange class << 8 | short-range class << 16
};

// The two rows of ordered class pair `index` of an overlap table.
__device__ __forceinline__ const OverlapRow* overlap_pair(const OverlapRow* table, int index) {
    return (const OverlapRow*) ((const float*) table+OVERLAP_PAIR_FLOATS*index);
}

// Pointers into LDS and the sizes of the tables of a launch.
struct PairTables {
    const float4* disp;      // numDisp^2 rows: c6, c8, c10, gate
    const OverlapRow* exch;  // 2 numExch^2 rows
    const OverlapRow* sr;    // 2 numSr^2 rows (== exch when the overlap is shared)
    int numDisp, numExch, numSr;
    float beta2;             // 2 * damping scale
    int shared;
    int highOrder;           // some row has non-zero p_4 .. p_8
    int singleBranch;        // every width pair takes one branch only (its two rows are equal): no row selection
};

// The row of a width pair that applies at distance r (the Taylor row where r*kTaylor < txLimit, else the closed
// form; with single-branch tables the first row, read without waiting for a selection).
__device__ __forceinline__ const OverlapRow& overlap_row(const PairTables& tb, const OverlapRow* rows, float r) {
    return (tb.singleBranch ? rows[0] : rows[(r*rows[0].kTaylor < rows[0].txLimit) ? 0 : 1]);
}

// S and dS/dr of a row: S = e1 (m1/r + sum_k p_k r^k) + e2 (q1/r + q0).
__device__ __forceinline__ void overlap_from_row(const PairTables& tb, const OverlapRow& w, float r, float invR, float& S, float& dS) {
    const float e1 = fast_exp2(-r*w.k1);
    const float e2 = fast_exp2(-r*w.k2);
    float P, dP = 0;
    if (tb.highOrder && __builtin_amdgcn_ballot_w64(w.high != 0.0f) != 0) {
        // some lane of the wave holds a degree-8 row
        P = w.ph[4];
#pragma unroll
        for (int k = 7; k >= 4; k--) {
            dP = dP*r + P;
            P = P*r + w.ph[k-4];
        }
        dP = dP*r + P;
        P = P*r + w.p3;
    }
    else {
        // every row of the wave is a closed form or an equal-width expansion: p_4 .. p_8 are zero
        P = w.p3;
    }
#pragma unroll
    for (int k = 2; k >= 0; k--) {
        dP = dP*r + P;
        P = P*r + w.p[k];
    }
    const float invR2 = invR*invR;
    const float P1 = P + w.m1*invR;
    const float dP1 = dP - w.m1*invR2;
    const float P2 = w.q1*invR + w.q0;
    const float dP2 = -w.q1*invR2;
    S = e1*P1 + e2*P2;
    dS = e1*(dP1 - w.n1*P1) + e2*(dP2 - w.n2*P2);
}

// Every table read is unconditional, so the compiler can issue them together: a class pair without a term has an
// all-zero overlap row (S = dS = 0 exactly, amplitude and norm 0) or a zero dispersion row, whose damping argument
// is replaced by a large constant so that every intermediate stays finite and the term contributes exactly 0.
// The fixed table form, for tables whose overlap rows take one branch each and are of degree 2 at most, whose
// dispersion rows are all live, and with at most one overlap per pair (exchange and short range share it, or only one
// of them is evaluated), independent of the particles' index order.  The kernel finds the pair's rows from offsets
// staged per particle (FixedPair), so the pair costs no class arithmetic; the overlap's energy terms come together
// as S (a1 + a2 S).
struct FixedPair {
    float4 disp;             // c6, c8, c10, gate
    const OverlapRow* row;
    float wsum;              // sum of the two dispersion widths
};

template <int TERMS>
__device__ __forceinline__ void pair_energy_fixed(float r, float invR, const FixedPair& fp, float beta2, float& energy, float& dEdR) {
    constexpr bool DISP = (TERMS & IXH_TERM_DISPERSION) != 0;
    constexpr bool OVERLAP = (TERMS & (IXH_TERM_EXCHANGE | IXH_TERM_SHORT_RANGE)) != 0;
    energy = 0;
    dEdR = 0;
    const float invR2 = invR*invR;
    if constexpr (DISP) {
        // E = -sum_n f_n e_n with e_n = C_n r^-n; dE/dr = sum_n e_n (n f_n / r - f_n' x'), f_n' = exp(-x) x^n/n!:
        // grouped as dE/dr = h/r - x' exp(-x) g with h = sum_n n f_n e_n and g = sum_n e_n x^n/n!
        const float dxdr = beta2*fast_rcp(fp.wsum);
        Damping<float> d;
        damping_main<float>(dxdr*r, d);
        damping_tail<float>(d);
        const float invR6 = invR2*invR2*invR2;
        const float invR8 = invR6*invR2;
        const float e6 = fp.disp.x*invR6;
        const float e8 = fp.disp.y*invR8;
        const float e10 = fp.disp.z*(invR8*invR2);
        const float p6 = d.f[0]*e6, p8 = d.f[1]*e8, p10 = d.f[2]*e10;
        const float h = 6.0f*p6 + 8.0f*p8 + 10.0f*p10;
        const float g = e6*d.t6 + e8*d.t8 + e10*d.t10;
        energy = -(p6 + p8 + p10);
        dEdR = h*invR - (dxdr*d.expX)*g;
    }
    if constexpr (OVERLAP) {
        const OverlapRow& w = *fp.row;
        const float e1 = fast_exp2(-r*w.k1);
        const float e2 = fast_exp2(-r*w.k2);
        // P = p0 + p1 r + p2 r^2, dP = p1 + 2 p2 r
        const float P1 = w.p[2]*r + w.p[1];
        const float dP = w.p[2]*r + P1;
        const float P = P1*r + w.p[0];
        const float Pa = P + w.m1*invR;
        const float dPa = dP - w.m1*invR2;
        const float Pb = w.q1*invR + w.q0;
        const float dPb = -w.q1*invR2;
        const float S = e1*Pa + e2*Pb;
        const float dS = e1*(dPa - w.n1*Pa) + e2*(dPb - w.n2*Pb);
        const float t = w.a1 + w.a2*S;
        energy += S*t;
        dEdR += dS*(t + w.a2*S);
    }
}

template <int TERMS>
__device__ __forceinline__ void pair_energy_table(float r, float invR, const TableParticle& pi, const TableParticle& pj, bool iFirst,
                                                  const PairTables& tb, float& energy, float& dEdR) {
    energy = 0;
    dEdR = 0;
    if (TERMS & IXH_TERM_DISPERSION) {
        const int di = pi.cls & 0xff, dj = pj.cls & 0xff;
        const float4 row = tb.disp[__umul24(di, tb.numDisp)+dj];
        const bool on = (row.w != 0.0f);
        const float dxdr = on ? tb.beta2*fast_rcp(pi.wd+pj.wd) : 0.0f;
        float f[3], df[3];
        tang_toennies<float>(on ? dxdr*r : 64.0f, f, df);
        const float invR2 = invR*invR;
        const float invR6 = invR2*invR2*invR2;
        const float invR8 = invR6*invR2;
        const float e6 = row.x*invR6;
        const float e8 = row.y*invR8;
        const float e10 = row.z*(invR8*invR2);
        energy -= f[0]*e6 + f[1]*e8 + f[2]*e10;
        dEdR -= e6*(df[0]*dxdr - 6.0f*f[0]*invR) + e8*(df[1]*dxdr - 8.0f*f[1]*invR) +
                e10*(df[2]*dxdr - 10.0f*f[2]*invR);
    }
    constexpr bool EXCH = (TERMS & IXH_TERM_EXCHANGE) != 0;
    constexpr bool SR = (TERMS & IXH_TERM_SHORT_RANGE) != 0;
    float sx = 0, dsx = 0;
    if (EXCH) {
        const int xi = (pi.cls >> 8) & 0xff, xj = (pj.cls >> 8) & 0xff;
        const OverlapRow& w = overlap_row(tb, overlap_pair(tb.exch, iFirst ? __umul24(xi, tb.numExch)+xj : __umul24(xj, tb.numExch)+xi), r);
        overlap_from_row(tb, w, r, invR, sx, dsx);
        const float amp = w.amplitude;
        energy += amp*sx;
        dEdR += amp*dsx;
    }
    if (SR) {
        const int si = (pi.cls >> 16) & 0xff, sj = (pj.cls >> 16) & 0xff;
        const OverlapRow& w = overlap_row(tb, overlap_pair(tb.sr, iFirst ? __umul24(si, tb.numSr)+sj : __umul24(sj, tb.numSr)+si), r);
        float ss, dss;
        if (EXCH && tb.shared) {
            ss = sx;
            dss = dsx;
        }
        else
            overlap_from_row(tb, w, r, invR, ss, dss);
        const float norm = w.norm;
        const float st = norm*ss;
        const float kk = w.kk;
        const float ll = w.ll;
        energy += (kk+ll*st)*st;
        dEdR += (kk+2.0f*ll*st)*norm*dss;
    }
}

}  // namespace ixh

#endif
