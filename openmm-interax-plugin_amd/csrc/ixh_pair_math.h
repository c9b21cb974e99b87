// Device functions of the InteraxFF pair energies: Tang-Toennies damped C6/C8/C10 dispersion, Slater-overlap
// exchange and the short-range correction, with their radial derivatives.  Templated over the arithmetic type:
// float for the single and mixed precision modes, double for the double precision mode.
//
// The mathematics is that of openmmapi/include/internal/InteraxPairTerms.h:167-337 (docs/SPEC.md:37-122); the
// evaluation is organised for the VALU of gfx950: reciprocals, square roots and exponentials use the hardware
// approximations (v_rcp_f32, v_rsq_f32, v_exp_f32) in FP32, both exponentials of a Slater overlap are shared
// between its closed form and its Taylor branch, and inverse widths are precomputed per particle.
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
    // coefficient of x^k of d/dx [exp(-x) sum V_k x^k] / exp(-x)
    __host__ __device__ static constexpr double d(int n, int k) {
        return (k+1 < 9 ? (k+1)*v(n, k+1) : 0.0) - v(n, k);
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
    Real l = (aboutA ? a : b);
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
        Real sumV = 0, sumD = 0;
#pragma unroll
        for (int n = OverlapTaylor::ORDER; n >= 0; n--) {
            if (n > scheme.order)
                continue;
            Real value = 0, dv = 0;
#pragma unroll
            for (int k = n+2; k >= 0; k--) {
                value = value*x + Real(OverlapTaylor::v(n, k));
                dv = dv*x + Real(OverlapTaylor::d(n, k));
            }
            sumV = sumV*t + value;
            sumD = sumD*t + dv;
        }
        Real prefactor = expL*invL*invL*invL*Real(INV_PI);
        taylorS = prefactor*sumV;
        taylorD = prefactor*invL*sumD;
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

// Tang-Toennies f_6, f_8, f_10 and df/dx (InteraxPairTerms.h:223-253).  Below x = (n+1)/2 the closed form
// 1 - exp(-x) sum_{k<=n} x^k/k! cancels and f_n = exp(-x) sum_{k>n} x^k/k! is used, a sum of positive terms.
template <typename Real>
__device__ __forceinline__ void tang_toennies(Real x, Real f[3], Real df[3]) {
    Real expX = fast_exp2(-x*Real(LOG2E));
    Real terms[11];
    terms[0] = 1;
#pragma unroll
    for (int k = 1; k <= 10; k++)
        terms[k] = terms[k-1]*x*Real(1.0/k);
    Real s6 = (((((terms[6]+terms[5])+terms[4])+terms[3])+terms[2])+terms[1])+terms[0];
    Real s8 = s6+terms[7]+terms[8];
    Real s10 = s8+terms[9]+terms[10];
    f[0] = 1-expX*s6;
    f[1] = 1-expX*s8;
    f[2] = 1-expX*s10;
    if (x < Real(5.5)) {
        Real term = terms[10], tail10 = 0;
#pragma unroll 6
        for (int k = 11; k <= 40; k++) {
            term *= x*Real(1.0/k);
            tail10 += term;
        }
        Real tail8 = tail10+terms[9]+terms[10];
        Real tail6 = tail8+terms[7]+terms[8];
        f[2] = expX*tail10;
        if (x < Real(4.5))
            f[1] = expX*tail8;
        if (x < Real(3.5))
            f[0] = expX*tail6;
    }
    df[0] = expX*terms[6];
    df[1] = expX*terms[8];
    df[2] = expX*terms[10];
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

}  // namespace ixh

#endif
