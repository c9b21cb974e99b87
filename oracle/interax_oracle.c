/*
 * FP64 CPU oracle of the InteraxFF pair and three-body forces -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the double precision Reference platform of the reference plugin.  It exists to check
 * the HIP product path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only as
 * the checker or the timed CPU baseline.  The product never links or calls it.
 *
 * Each function names the reference file:line it follows.  Paths are relative to the reference repository root.
 * What pins each function (golden vectors, closed forms, library routines, brute force) is listed in the tests
 * under tests/test_oracle_*.py.
 */
#include "interax_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------------------------------------------ */
/* Slater overlap                                                                                                */
/* ------------------------------------------------------------------------------------------------------------ */

/*
 * Taylor coefficients of pi l^3 exp(x) S(x l; l, l (1 + t)) = sum_n t^n sum_k V[n][k] x^k, derived symbolically
 * from the closed form by tests/golden/derive_overlap_taylor.py (exact rationals).  The same table is
 * InteraxPairTerms.h:96-104 (a mathematical table: it must be alike).  The derivative table follows from
 * d/dx [exp(-x) sum V_k x^k] = exp(-x) sum ((k+1) V_{k+1} - V_k) x^k, as InteraxPairTerms.h:92 states.
 */
#define ORC_TAYLOR_ORDER 6
#define ORC_TAYLOR_TERMS 9
static const double OVERLAP_V[ORC_TAYLOR_ORDER+1][ORC_TAYLOR_TERMS] = {
    {1.0/64, 1.0/64, 1.0/192},
    {-3.0/128, -3.0/128, -1.0/192, 1.0/384},
    {3.0/128, 3.0/128, 1.0/768, -5.0/768, 1.0/1280},
    {-5.0/256, -5.0/256, 1.0/256, 1.0/96, -11.0/3840, 1.0/5760},
    {15.0/1024, 15.0/1024, -9.0/1024, -7.0/512, 33.0/5120, -19.0/23040, 1.0/32256},
    {-21.0/2048, -21.0/2048, 13.0/1024, 33.0/2048, -59.0/5120, 107.0/46080, -29.0/161280, 1.0/215040},
    {7.0/1024, 7.0/1024, -191.0/12288, -73.0/4096, 369.0/20480, -233.0/46080, 389.0/645120, -41.0/1290240, 1.0/1658880}
};

static double overlap_d(int n, int k) {
    double next = (k+1 < ORC_TAYLOR_TERMS ? (k+1)*OVERLAP_V[n][k+1] : 0.0);
    return next - OVERLAP_V[n][k];
}

/* InteraxPairTerms.h:139-147 */
void orc_scheme_accurate(int double_precision, orc_taylor_scheme* s) {
    s->order = 6;
    s->about_smaller_width = 1;
    s->max_delta = 0.0;
    s->max_t = (double_precision ? 0.02 : 0.2);
    s->max_tx = (double_precision ? 0.3 : 1.6);
}

/* InteraxPairTerms.h:151-159 (0.01 bohr) */
void orc_scheme_order2(orc_taylor_scheme* s) {
    s->order = 2;
    s->about_smaller_width = 0;
    s->max_delta = 0.01*0.0529177210903;
    s->max_t = 0.0;
    s->max_tx = HUGE_VAL;
}

/* InteraxPairTerms.h:167-216.  The reference evaluates both branches and blends; only the selected one is kept. */
void orc_slater_overlap(double r, double a, double b, const orc_taylor_scheme* s, double* overlap, double* deriv) {
    double l = (s->about_smaller_width ? fmin(a, b) : a);
    double invL = 1.0/l;
    double delta = a+b-l-l;
    double t = delta*invL;
    double x = r*invL;
    double absT = fabs(t);
    int nearAbsolute = (fabs(delta) < s->max_delta);
    int nearRelative = (absT < s->max_t) && (absT*x < s->max_tx);
    if (nearAbsolute || nearRelative) {
        double sumValue = 0, sumDeriv = 0;
        for (int n = s->order; n >= 0; n--) {
            double value = 0, dv = 0;
            for (int k = ORC_TAYLOR_TERMS-1; k >= 0; k--) {
                value = value*x + OVERLAP_V[n][k];
                dv = dv*x + overlap_d(n, k);
            }
            sumValue = sumValue*t + value;
            sumDeriv = sumDeriv*t + dv;
        }
        double prefactor = exp(-x)*invL*invL*invL/M_PI;
        *overlap = prefactor*sumValue;
        *deriv = prefactor*invL*sumDeriv;
        return;
    }
    /* closed form S = F/(8 pi r) */
    double aa = a*a, bb = b*b;
    double d = bb-aa;
    double invD2 = 1/(d*d);
    double expA = exp(-r/a), expB = exp(-r/b);
    double c = 4*aa*bb*invD2/d;
    double f = c*(expA-expB) + r*invD2*(a*expA+b*expB);
    double dfdr = c*(expB/b-expA/a) + invD2*(a*expA+b*expB-r*(expA+expB));
    double invR = 1/r;
    *overlap = f*invR/(8*M_PI);
    *deriv = (r*dfdr-f)*invR*invR/(8*M_PI);
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Pair terms                                                                                                    */
/* ------------------------------------------------------------------------------------------------------------ */

/* InteraxPairTerms.h:223-253 (series tail to k = 40, switch at x = (n+1)/2) */
void orc_tang_toennies(double x, double damping[3], double dd[3]) {
    const int maxTerm = 40;
    double expX = exp(-x);
    double term = 1, partial = 1;
    double terms[11], partials[11];
    terms[0] = term;
    partials[0] = partial;
    for (int k = 1; k <= 10; k++) {
        term = term*x*(1.0/k);
        partial = partial+term;
        terms[k] = term;
        partials[k] = partial;
    }
    double tail = 0;
    for (int k = 11; k <= maxTerm; k++) {
        term = term*x*(1.0/k);
        tail = tail+term;
    }
    for (int i = 0; i < 3; i++) {
        int n = 6+2*i;
        double series = tail;
        for (int k = n+1; k <= 10; k++)
            series = series+terms[k];
        damping[i] = (x < 0.5*(n+1) ? expX*series : 1-expX*partials[n]);
        dd[i] = expX*terms[n];
    }
}

/* InteraxPairTerms.h:267-293.  p = (m1, m2, m3, alpha, width). */
void orc_dispersion_pair(double r, const double* pi, const double* pj, double dampingScale, double* energy, double* dEdR) {
    double m1i = pi[0], m2i = pi[1], m3i = pi[2], ai = pi[3], wi = pi[4];
    double m1j = pj[0], m2j = pj[1], m3j = pj[2], aj = pj[3], wj = pj[4];
    double prefactor = ai*aj/fmax(ai*m1j+aj*m1i, 1e-30);
    double c[3];
    c[0] = prefactor*m1i*m1j;
    c[1] = 1.5*prefactor*(m1i*m2j+m2i*m1j);
    c[2] = prefactor*(2*(m1i*m3j+m3i*m1j)+4.2*m2i*m2j);
    double dxdr = 2*dampingScale/(wi+wj);
    double f[3], df[3];
    orc_tang_toennies(dxdr*r, f, df);
    double invR = 1/r;
    double invR2 = invR*invR;
    double invRn = invR2*invR2*invR2;
    double e = 0, d = 0;
    for (int i = 0; i < 3; i++) {
        double n = 6+2*i;
        e = e - f[i]*c[i]*invRn;
        d = d - c[i]*invRn*(df[i]*dxdr - n*f[i]*invR);
        invRn = invRn*invR2;
    }
    *energy = e;
    *dEdR = d;
}

/* InteraxPairTerms.h:305-312.  p = (amplitude, width). */
void orc_exchange_pair(double r, const double* pi, const double* pj, const orc_taylor_scheme* s, double* energy, double* dEdR) {
    double S, dS;
    orc_slater_overlap(r, pi[1], pj[1], s, &S, &dS);
    double amplitude = pi[0]*pj[0];
    *energy = amplitude*S;
    *dEdR = amplitude*dS;
}

/* InteraxPairTerms.h:325-337.  p = (k, l, width). */
void orc_short_range_pair(double r, const double* pi, const double* pj, const orc_taylor_scheme* s, double* energy, double* dEdR) {
    double S, dS;
    orc_slater_overlap(r, pi[2], pj[2], s, &S, &dS);
    double widthSum = pi[2]+pj[2];
    double normalization = (8*M_PI)*widthSum*widthSum*widthSum;
    double St = normalization*S;
    double k = 0.5*(pi[0]+pj[0]);
    double l = 0.5*(pi[1]+pj[1]);
    *energy = (k+l*St)*St;
    *dEdR = (k+2*l*St)*normalization*dS;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Geometry                                                                                                      */
/* ------------------------------------------------------------------------------------------------------------ */

/*
 * Minimum image displacement to - from, reduced along c, b, a in turn: OpenMM's ReferenceForce::getDeltaRPeriodic
 * as called at ReferenceInteraxPairIxn.h:93-96, and the same formula as ReferenceInteraxThreeBodyIxn.cpp:74-82.
 */
static void delta_vec(const double* from, const double* to, const double* box, int periodic, double d[3]) {
    d[0] = to[0]-from[0];
    d[1] = to[1]-from[1];
    d[2] = to[2]-from[2];
    if (periodic) {
        double s2 = floor(d[2]/box[8]+0.5);
        d[0] -= box[6]*s2; d[1] -= box[7]*s2; d[2] -= box[8]*s2;
        double s1 = floor(d[1]/box[4]+0.5);
        d[0] -= box[3]*s1; d[1] -= box[4]*s1; d[2] -= box[5]*s1;
        double s0 = floor(d[0]/box[0]+0.5);
        d[0] -= box[0]*s0; d[1] -= box[1]*s0; d[2] -= box[2]*s0;
    }
}

static double dot3(const double* a, const double* b) {
    return a[0]*b[0] + a[1]*b[1] + a[2]*b[2];
}

/*
 * Candidate pairs (i < j) of a point set: every pair whose distance can be below `cutoff`.  A cell grid is used for
 * rectangular periodic boxes with at least 3 cells of side >= cutoff along every axis and for large nonperiodic
 * sets; otherwise all pairs.  The predicate itself is always evaluated by the caller with delta_vec.
 */
typedef void (*pair_visitor)(int i, int j, void* data);

static int is_rectangular(const double* box) {
    return box[1] == 0 && box[2] == 0 && box[3] == 0 && box[5] == 0 && box[6] == 0 && box[7] == 0;
}

static void for_each_candidate_pair(int n, const double* pos, const double* box, int periodic, double cutoff,
                                    pair_visitor visit, void* data) {
    int use_cells = 0;
    int ncell[3] = {1, 1, 1};
    double lo[3] = {0, 0, 0}, size[3] = {1, 1, 1};
    if (n > 256 && cutoff > 0) {
        if (periodic) {
            if (is_rectangular(box)) {
                use_cells = 1;
                for (int k = 0; k < 3; k++) {
                    double L = box[4*k];
                    ncell[k] = (int) floor(L/cutoff);
                    if (ncell[k] < 3)
                        use_cells = 0;
                    if (ncell[k] > 512)
                        ncell[k] = 512;
                    size[k] = L/ncell[k];
                }
            }
        }
        else {
            double hi[3];
            for (int k = 0; k < 3; k++) {
                lo[k] = hi[k] = pos[k];
            }
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++) {
                    lo[k] = fmin(lo[k], pos[3*i+k]);
                    hi[k] = fmax(hi[k], pos[3*i+k]);
                }
            use_cells = 1;
            for (int k = 0; k < 3; k++) {
                ncell[k] = (int) floor((hi[k]-lo[k])/cutoff);
                if (ncell[k] < 1)
                    ncell[k] = 1;
                if (ncell[k] > 512)
                    ncell[k] = 512;
                size[k] = (hi[k]-lo[k])/ncell[k];
                if (size[k] < cutoff)
                    size[k] = cutoff;
            }
        }
    }
    if (!use_cells) {
        for (int i = 0; i < n; i++)
            for (int j = i+1; j < n; j++)
                visit(i, j, data);
        return;
    }
    long long total = (long long) ncell[0]*ncell[1]*ncell[2];
    int* head = (int*) malloc(sizeof(int)*total);
    int* next = (int*) malloc(sizeof(int)*n);
    int* cellOf = (int*) malloc(sizeof(int)*n);
    for (long long c = 0; c < total; c++)
        head[c] = -1;
    for (int i = n-1; i >= 0; i--) {
        int idx[3];
        for (int k = 0; k < 3; k++) {
            double p = pos[3*i+k];
            double u = (periodic ? p/size[k] : (p-lo[k])/size[k]);
            int c = (int) floor(u);
            if (periodic) {
                c %= ncell[k];
                if (c < 0)
                    c += ncell[k];
            }
            else {
                if (c < 0) c = 0;
                if (c >= ncell[k]) c = ncell[k]-1;
            }
            idx[k] = c;
        }
        int cell = (idx[2]*ncell[1]+idx[1])*ncell[0]+idx[0];
        cellOf[i] = cell;
        next[i] = head[cell];
        head[cell] = i;
    }
    for (int i = 0; i < n; i++) {
        int cell = cellOf[i];
        int cx = cell%ncell[0], cy = (cell/ncell[0])%ncell[1], cz = cell/(ncell[0]*ncell[1]);
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int x = cx+dx, y = cy+dy, z = cz+dz;
                    if (periodic) {
                        x = (x+ncell[0])%ncell[0];
                        y = (y+ncell[1])%ncell[1];
                        z = (z+ncell[2])%ncell[2];
                    }
                    else if (x < 0 || y < 0 || z < 0 || x >= ncell[0] || y >= ncell[1] || z >= ncell[2])
                        continue;
                    int other = (z*ncell[1]+y)*ncell[0]+x;
                    for (int j = head[other]; j >= 0; j = next[j])
                        if (j > i)
                            visit(i, j, data);
                }
    }
    free(head);
    free(next);
    free(cellOf);
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Pair force loop                                                                                               */
/* ------------------------------------------------------------------------------------------------------------ */

/* Exclusions as per-particle sorted lists (the sets of ReferenceInteraxPairIxn.cpp:43-49). */
typedef struct {
    int* offsets;
    int* partners;
} exclusion_table;

static int cmp_int(const void* a, const void* b) {
    int x = *(const int*) a, y = *(const int*) b;
    return (x > y) - (x < y);
}

static void build_exclusions(const orc_pair_force* f, exclusion_table* t) {
    int n = f->num_particles;
    t->offsets = (int*) calloc((size_t) n+1, sizeof(int));
    t->partners = (int*) malloc(sizeof(int)*(2*f->num_exclusions+1));
    for (int e = 0; e < f->num_exclusions; e++) {
        t->offsets[f->exclusions[2*e]+1]++;
        t->offsets[f->exclusions[2*e+1]+1]++;
    }
    for (int i = 0; i < n; i++)
        t->offsets[i+1] += t->offsets[i];
    int* fill = (int*) calloc((size_t) n+1, sizeof(int));
    for (int e = 0; e < f->num_exclusions; e++) {
        int p1 = f->exclusions[2*e], p2 = f->exclusions[2*e+1];
        t->partners[t->offsets[p1]+fill[p1]++] = p2;
        t->partners[t->offsets[p2]+fill[p2]++] = p1;
    }
    for (int i = 0; i < n; i++)
        qsort(t->partners+t->offsets[i], t->offsets[i+1]-t->offsets[i], sizeof(int), cmp_int);
    free(fill);
}

static int is_excluded(const exclusion_table* t, int i, int j) {
    int lo = t->offsets[i], hi = t->offsets[i+1];
    while (lo < hi) {
        int mid = (lo+hi)/2;
        int v = t->partners[mid];
        if (v == j)
            return 1;
        if (v < j)
            lo = mid+1;
        else
            hi = mid;
    }
    return 0;
}

static int row_size(int kind) {
    return (kind == ORC_DISPERSION ? 5 : kind == ORC_EXCHANGE ? 2 : 3);
}

typedef struct {
    const orc_pair_force* f;
    const double* pos;
    const double* box;
    exclusion_table excl;
    orc_taylor_scheme scheme;
    double lambda;
    double* forces;
    double energy;
    double abs_energy;     /* sum of |pair energy| before switching: the scale of rounding errors */
    long long count;
    int count_only;
} pair_loop_state;

/*
 * The pair energy callbacks of ReferenceInteraxKernels.cpp:84-89 (dispersion, gate alpha_i alpha_j != 0),
 * :130-140 (exchange) and :173-183 (short range; both gate width_i width_j != 0 and order the pair by index).
 */
static int pair_energy(const pair_loop_state* s, int i, int j, double r, double* e, double* dEdR) {
    const orc_pair_force* f = s->f;
    int rs = row_size(f->kind);
    if (f->kind == ORC_DISPERSION) {
        const double* p1 = f->params+rs*i;
        const double* p2 = f->params+rs*j;
        if (p1[3]*p2[3] == 0)
            return 0;
        orc_dispersion_pair(r, p1, p2, f->damping_scale, e, dEdR);
        return 1;
    }
    int lo = (i < j ? i : j), hi = (i < j ? j : i);
    const double* p1 = f->params+rs*lo;
    const double* p2 = f->params+rs*hi;
    if (f->kind == ORC_EXCHANGE) {
        if (p1[1]*p2[1] == 0)
            return 0;
        orc_exchange_pair(r, p1, p2, &s->scheme, e, dEdR);
    }
    else {
        if (p1[2]*p2[2] == 0)
            return 0;
        orc_short_range_pair(r, p1, p2, &s->scheme, e, dEdR);
    }
    return 1;
}

/* The body of the loop of ReferenceInteraxPairIxn.h:60-97 for one pair. */
static void visit_pair(int i, int j, void* data) {
    pair_loop_state* s = (pair_loop_state*) data;
    const orc_pair_force* f = s->f;
    if (is_excluded(&s->excl, i, j))
        return;
    int periodic = (f->method == ORC_CUTOFF_PERIODIC);
    double d[3];
    delta_vec(s->pos+3*i, s->pos+3*j, s->box, periodic, d);
    double r = sqrt(dot3(d, d));
    if (f->method != ORC_NO_CUTOFF && r >= f->cutoff)
        return;
    int alchemicalPair = (f->alchemical != NULL && (f->alchemical[i] != 0) != (f->alchemical[j] != 0));
    double rEff = r;
    if (alchemicalPair && f->softcore_radius > 0) {
        double R3 = f->softcore_radius*f->softcore_radius*f->softcore_radius;
        rEff = pow(r*r*r*r*r*r + (1-s->lambda)*R3*R3, 1.0/6.0);
    }
    double e, dEdR;
    if (!pair_energy(s, i, j, rEff, &e, &dEdR))
        return;
    s->count++;
    if (s->count_only)
        return;
    if (rEff != r)
        dEdR *= pow(r/rEff, 5);
    s->abs_energy += fabs((alchemicalPair ? s->lambda : 1.0)*e);
    int useSwitch = (f->method != ORC_NO_CUTOFF && f->use_switch);
    if (useSwitch && r > f->switching_distance) {
        double width = f->cutoff-f->switching_distance;
        double t = (r-f->switching_distance)/width;
        double sv = 1+t*t*t*(-10+t*(15-t*6));
        double sd = t*t*(-30+t*(60-t*30))/width;
        dEdR = dEdR*sv + e*sd;
        e *= sv;
    }
    double scale = (alchemicalPair ? s->lambda : 1.0);
    s->energy += scale*e;
    double forceScale = scale*dEdR/r;
    for (int k = 0; k < 3; k++) {
        s->forces[3*i+k] += d[k]*forceScale;
        s->forces[3*j+k] -= d[k]*forceScale;
    }
}

static int any_alchemical(const unsigned char* flags, int n) {
    if (flags == NULL)
        return 0;
    for (int i = 0; i < n; i++)
        if (flags[i])
            return 1;
    return 0;
}

static int run_pair_loop(const orc_pair_force* f, const double* positions, const double* box, double lambda,
                         double* forces, double* energy, long long* count, int count_only, double* abs_energy) {
    if (f == NULL || positions == NULL || f->num_particles < 0 || f->kind < 0 || f->kind > 2)
        return -1;
    if (f->method == ORC_CUTOFF_PERIODIC && box == NULL)
        return -1;
    pair_loop_state s;
    memset(&s, 0, sizeof(s));
    s.f = f;
    s.pos = positions;
    s.box = box;
    /* ReferenceInteraxKernels.cpp:63-65: the context parameter is read only if a particle is alchemical */
    s.lambda = (any_alchemical(f->alchemical, f->num_particles) ? lambda : 1.0);
    s.forces = forces;
    s.count_only = count_only;
    /* InteraxPairForceImpl.h:155-160 */
    if (f->overlap_method == 0)
        orc_scheme_accurate(f->double_precision, &s.scheme);
    else
        orc_scheme_order2(&s.scheme);
    build_exclusions(f, &s.excl);
    if (f->method == ORC_NO_CUTOFF) {
        /* ReferenceInteraxPairIxn.cpp:56-60: all non-excluded pairs, in order */
        for (int i = 0; i < f->num_particles; i++)
            for (int j = i+1; j < f->num_particles; j++)
                visit_pair(i, j, &s);
    }
    else
        for_each_candidate_pair(f->num_particles, positions, box, f->method == ORC_CUTOFF_PERIODIC, f->cutoff, visit_pair, &s);
    free(s.excl.offsets);
    free(s.excl.partners);
    if (count != NULL)
        *count = s.count;
    if (!count_only && f->kind == ORC_DISPERSION) {
        double coef, alch;
        orc_long_range_correction(f, &coef, &alch);
        /* ReferenceInteraxKernels.cpp:93-94 */
        if (coef != 0 || alch != 0)
            s.energy += (coef+s.lambda*alch)/(box[0]*box[4]*box[8]);
    }
    if (energy != NULL)
        *energy = s.energy;
    if (abs_energy != NULL)
        *abs_energy = s.abs_energy;
    return 0;
}

int orc_pair_compute(const orc_pair_force* f, const double* positions, const double* box, double lambda,
                     double* forces, double* energy) {
    return run_pair_loop(f, positions, box, lambda, forces, energy, NULL, 0, NULL);
}

int orc_pair_compute_scale(const orc_pair_force* f, const double* positions, const double* box, double lambda,
                           double* forces, double* energy, double* abs_energy) {
    return run_pair_loop(f, positions, box, lambda, forces, energy, NULL, 0, abs_energy);
}

long long orc_pair_count(const orc_pair_force* f, const double* positions, const double* box) {
    long long count = 0;
    if (run_pair_loop(f, positions, box, 1.0, NULL, NULL, &count, 1, NULL) != 0)
        return -1;
    return count;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Long range correction                                                                                         */
/* ------------------------------------------------------------------------------------------------------------ */

/* InteraxDispersionForceImpl.cpp:49-73: Newton iteration for the roots of P_n from the Chebyshev estimate. */
void orc_gauss_legendre(int n, double* nodes, double* weights) {
    for (int i = 0; i < n; i++) {
        double x = cos(M_PI*(i+0.75)/(n+0.5));
        double derivative = 0;
        for (int iteration = 0; iteration < 100; iteration++) {
            double p0 = 1, p1 = x;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2*k-1)*x*p1-(k-1)*p0)/k;
                p0 = p1;
                p1 = p2;
            }
            derivative = n*(x*p1-p0)/(x*x-1);
            double step = p1/derivative;
            x -= step;
            if (fabs(step) < 1e-15)
                break;
        }
        nodes[i] = 0.5*(1-x);
        weights[i] = 1/((1-x*x)*derivative*derivative);
    }
}

typedef struct {
    double p[5];
    int alchemical;
} particle_class;

static int cmp_class(const void* a, const void* b) {
    const particle_class* x = (const particle_class*) a;
    const particle_class* y = (const particle_class*) b;
    for (int k = 0; k < 5; k++) {
        if (x->p[k] < y->p[k])
            return -1;
        if (x->p[k] > y->p[k])
            return 1;
    }
    return (x->alchemical > y->alchemical) - (x->alchemical < y->alchemical);
}

/* InteraxDispersionForceImpl.cpp:86-149 */
void orc_long_range_correction(const orc_pair_force* f, double* coefficient, double* alchemicalCoefficient) {
    *coefficient = 0;
    *alchemicalCoefficient = 0;
    if (f->kind != ORC_DISPERSION || f->method != ORC_CUTOFF_PERIODIC || !f->use_long_range_correction)
        return;
    int n = f->num_particles;
    /* Classes of identical parameters, ordered as the keys of a std::map<tuple> (lexicographic). */
    particle_class* all = (particle_class*) malloc(sizeof(particle_class)*(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        memcpy(all[i].p, f->params+5*i, 5*sizeof(double));
        all[i].alchemical = (f->alchemical != NULL && f->alchemical[i] != 0);
    }
    qsort(all, n, sizeof(particle_class), cmp_class);
    particle_class* classes = (particle_class*) malloc(sizeof(particle_class)*(n > 0 ? n : 1));
    double* counts = (double*) malloc(sizeof(double)*(n > 0 ? n : 1));
    int numClasses = 0;
    for (int i = 0; i < n; i++) {
        if (numClasses > 0 && cmp_class(&classes[numClasses-1], &all[i]) == 0)
            counts[numClasses-1] += 1;
        else {
            classes[numClasses] = all[i];
            counts[numClasses] = 1;
            numClasses++;
        }
    }
    const int numPoints = 32;
    double nodes[32], weights[32];
    orc_gauss_legendre(numPoints, nodes, weights);
    double cutoff = f->cutoff;
    int useSwitch = f->use_switch;
    double switchDistance = (useSwitch ? f->switching_distance : cutoff);
    double switchWidth = cutoff-switchDistance;
    double sum = 0, alchemicalSum = 0;
    for (int i = 0; i < numClasses; i++)
        for (int j = i; j < numClasses; j++) {
            const particle_class* c1 = &classes[i];
            const particle_class* c2 = &classes[j];
            if (c1->p[3]*c2->p[3] == 0)
                continue;
            double integral = 0;
            for (int k = 0; k < numPoints; k++) {
                double u = nodes[k];
                double e, d;
                orc_dispersion_pair(cutoff/u, c1->p, c2->p, f->damping_scale, &e, &d);
                integral += weights[k]*e*cutoff*cutoff*cutoff/(u*u*u*u);
                if (useSwitch) {
                    double x = nodes[k];
                    double r = switchDistance+x*switchWidth;
                    double oneMinusSwitch = x*x*x*(10+x*(-15+x*6));
                    orc_dispersion_pair(r, c1->p, c2->p, f->damping_scale, &e, &d);
                    integral += weights[k]*switchWidth*oneMinusSwitch*e*r*r;
                }
            }
            double count1 = counts[i], count2 = counts[j];
            double count = (i == j ? 0.5*count1*(count1+1) : count1*count2);
            if (c1->alchemical || c2->alchemical)
                alchemicalSum += count*integral;
            else
                sum += count*integral;
        }
    double N = n;
    double normalization = 2*M_PI*N*N/(0.5*N*(N+1));
    *coefficient = normalization*sum;
    *alchemicalCoefficient = normalization*alchemicalSum;
    free(all);
    free(classes);
    free(counts);
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Three-body                                                                                                    */
/* ------------------------------------------------------------------------------------------------------------ */

/* InteraxThreeBodyTerms.h:51-65 */
void orc_tang_toennies3(double x, double* damping, double* deriv) {
    double expX = exp(-x);
    double term3 = x*x*x/6;
    *deriv = expX*term3;
    if (x < 2) {
        double term = term3, tail = 0;
        for (int k = 4; k <= 30; k++) {
            term *= x/k;
            tail += term;
        }
        *damping = expX*tail;
    }
    else
        *damping = 1-expX*(1+x+x*x/2+term3);
}

/* InteraxThreeBodyTerms.h:71-77 */
static void centroid_damping(double r, double invRadius, double* damping, double* deriv) {
    double x = r*invRadius;
    double x6 = x*x*x;
    x6 *= x6;
    *damping = -expm1(-x6);
    *deriv = 6*x6*exp(-x6)/r;
}

/* InteraxThreeBodyForceImpl.h:69-72 */
double orc_c9(double a1, double w1, double a2, double w2, double a3, double w3) {
    return 1.5*a1*a2*a3*w1*w2*w3*(w1+w2+w3)/((w1+w2)*(w2+w3)*(w3+w1));
}

/* InteraxThreeBodyTerms.h:97-146 */
double orc_triple(const double a[3], const double b[3], const double c[3], double c9, const double* rates,
                  double fa[3], double fb[3], double fc[3], double dampingRadius) {
    double u = dot3(a, a), v = dot3(b, b), w = dot3(c, c);
    double p = dot3(a, b), q = dot3(a, c), s = dot3(b, c);
    double invP = 1/(u*v*w);
    double invP32 = invP*sqrt(invP);
    double invP52 = invP32*invP;
    double pqs = p*q*s;
    double energy = c9*(invP32 - 3*pqs*invP52);
    double radial = c9*(-3*invP32 + 15*pqs*invP52);
    double cross = -3*c9*invP52;
    double dEda[3], dEdb[3], dEdc[3];
    for (int k = 0; k < 3; k++) {
        dEda[k] = a[k]*(radial/u) + (b[k]*(q*s) + c[k]*(p*s))*cross;
        dEdb[k] = b[k]*(radial/v) + (a[k]*(q*s) + c[k]*(p*q))*cross;
        dEdc[k] = c[k]*(radial/w) + (a[k]*(p*s) + b[k]*(p*q))*cross;
    }
    if (rates != NULL || dampingRadius > 0) {
        double lengths[3] = {sqrt(u), sqrt(v), sqrt(w)};
        double f[3], dfdr[3];
        for (int i = 0; i < 3; i++) {
            if (rates != NULL) {
                orc_tang_toennies3(rates[i]*lengths[i], &f[i], &dfdr[i]);
                dfdr[i] *= rates[i];
            }
            else {
                f[i] = 1;
                dfdr[i] = 0;
            }
            if (dampingRadius > 0) {
                double g, dg;
                centroid_damping(lengths[i], 1/dampingRadius, &g, &dg);
                dfdr[i] = dfdr[i]*g + f[i]*dg;
                f[i] *= g;
            }
        }
        double damping = f[0]*f[1]*f[2];
        double sa = energy*dfdr[0]*f[1]*f[2]/lengths[0];
        double sb = energy*f[0]*dfdr[1]*f[2]/lengths[1];
        double sc = energy*f[0]*f[1]*dfdr[2]/lengths[2];
        for (int k = 0; k < 3; k++) {
            dEda[k] = dEda[k]*damping + a[k]*sa;
            dEdb[k] = dEdb[k]*damping + b[k]*sb;
            dEdc[k] = dEdc[k]*damping + c[k]*sc;
        }
        energy *= damping;
    }
    for (int k = 0; k < 3; k++) {
        fa[k] += dEda[k] + dEdb[k];
        fb[k] += dEdc[k] - dEda[k];
        fc[k] -= dEdb[k] + dEdc[k];
    }
    return energy;
}

typedef struct {
    const orc_threebody_force* f;
    const double* sites;
    const double* box;
    const int* molecule;
    double cutoff2;
    int* count;      /* neighbors above each site */
    int** lists;
    int* capacity;
} tb_neighbor_state;

static void tb_visit(int i, int j, void* data) {
    tb_neighbor_state* s = (tb_neighbor_state*) data;
    if (s->molecule[i] == s->molecule[j])
        return;
    double d[3];
    delta_vec(s->sites+3*i, s->sites+3*j, s->box, s->f->periodic, d);
    if (!(dot3(d, d) < s->cutoff2))
        return;
    if (s->count[i] == s->capacity[i]) {
        s->capacity[i] = 2*s->capacity[i]+8;
        s->lists[i] = (int*) realloc(s->lists[i], sizeof(int)*s->capacity[i]);
    }
    s->lists[i][s->count[i]++] = j;
}

static int run_threebody(const orc_threebody_force* f, const double* positions, const double* box, double lambda,
                         double* forces, double* energy_out, long long* count_out, double* abs_out) {
    if (f == NULL || positions == NULL || (f->periodic && box == NULL))
        return -1;
    int n = f->num_sites;
    /* ReferenceInteraxThreeBodyIxn.cpp:87-93: unweighted centroids (OpenMM's Vec3 divides by multiplying with
       the reciprocal) */
    double* sites = (double*) calloc(3*(size_t) (n > 0 ? n : 1), sizeof(double));
    int* molecule = (int*) malloc(sizeof(int)*(n > 0 ? n : 1));
    int anyAlchemical = 0;
    for (int i = 0; i < n; i++) {
        double s[3] = {0, 0, 0};
        for (int p = f->site_offsets[i]; p < f->site_offsets[i+1]; p++)
            for (int k = 0; k < 3; k++)
                s[k] += positions[3*f->site_particles[p]+k];
        double scale = 1.0/(f->site_offsets[i+1]-f->site_offsets[i]);
        for (int k = 0; k < 3; k++)
            sites[3*i+k] = s[k]*scale;
        /* ReferenceInteraxThreeBodyIxn.cpp:48-49 */
        molecule[i] = (f->molecule[i] < 0 ? -1-i : f->molecule[i]);
        if (f->alchemical != NULL && f->alchemical[i])
            anyAlchemical = 1;
    }
    double lam = (anyAlchemical ? lambda : 1.0);
    /* ReferenceInteraxThreeBodyIxn.cpp:97-103: neighbors of different molecules with |d|^2 < cutoff^2, kept as
       ascending lists of the sites above each site */
    tb_neighbor_state ns;
    ns.f = f;
    ns.sites = sites;
    ns.box = box;
    ns.molecule = molecule;
    ns.cutoff2 = f->cutoff*f->cutoff;
    ns.count = (int*) calloc(n > 0 ? n : 1, sizeof(int));
    ns.capacity = (int*) calloc(n > 0 ? n : 1, sizeof(int));
    ns.lists = (int**) calloc(n > 0 ? n : 1, sizeof(int*));
    for_each_candidate_pair(n, sites, box, f->periodic, f->cutoff, tb_visit, &ns);
    for (int i = 0; i < n; i++)
        qsort(ns.lists[i], ns.count[i], sizeof(int), cmp_int);
    /* ReferenceInteraxThreeBodyIxn.cpp:107-125: i < j < k, all three pairs neighbors, in lexicographic order */
    double* siteForces = (double*) calloc(3*(size_t) (n > 0 ? n : 1), sizeof(double));
    double energy = 0, absEnergy = 0;
    long long count = 0;
    int useDamping = (f->damping_method == 1);
    for (int i = 0; i < n; i++) {
        for (int a = 0; a < ns.count[i]; a++) {
            int j = ns.lists[i][a];
            for (int b = a+1; b < ns.count[i]; b++) {
                int k = ns.lists[i][b];
                /* is k a neighbor of j?  Same predicate as the list. */
                if (molecule[j] == molecule[k])
                    continue;
                double djk[3];
                delta_vec(sites+3*j, sites+3*k, box, f->periodic, djk);
                if (!(dot3(djk, djk) < ns.cutoff2))
                    continue;
                double c9 = f->scale_factor*orc_c9(f->alpha[i], f->frequency[i], f->alpha[j], f->frequency[j], f->alpha[k], f->frequency[k]);
                if (f->alchemical != NULL && (f->alchemical[i] || f->alchemical[j] || f->alchemical[k]))
                    c9 *= lam;
                if (c9 == 0)
                    continue;
                count++;
                if (count_out != NULL && forces == NULL)
                    continue;
                double rates[3] = {2*f->damping_scale/(f->width[i]+f->width[j]), 2*f->damping_scale/(f->width[i]+f->width[k]),
                                   2*f->damping_scale/(f->width[j]+f->width[k])};
                double dij[3], dik[3];
                delta_vec(sites+3*i, sites+3*j, box, f->periodic, dij);
                delta_vec(sites+3*i, sites+3*k, box, f->periodic, dik);
                double e = orc_triple(dij, dik, djk, c9, useDamping ? rates : NULL, siteForces+3*i, siteForces+3*j,
                                      siteForces+3*k, f->centroid_damping_radius);
                energy += e;
                absEnergy += fabs(e);
            }
        }
    }
    /* ReferenceInteraxThreeBodyIxn.cpp:129-133 */
    if (forces != NULL)
        for (int i = 0; i < n; i++) {
            double scale = 1.0/(f->site_offsets[i+1]-f->site_offsets[i]);
            for (int p = f->site_offsets[i]; p < f->site_offsets[i+1]; p++)
                for (int k = 0; k < 3; k++)
                    forces[3*f->site_particles[p]+k] += siteForces[3*i+k]*scale;
        }
    for (int i = 0; i < n; i++)
        free(ns.lists[i]);
    free(ns.lists);
    free(ns.count);
    free(ns.capacity);
    free(siteForces);
    free(sites);
    free(molecule);
    if (energy_out != NULL)
        *energy_out = energy;
    if (count_out != NULL)
        *count_out = count;
    if (abs_out != NULL)
        *abs_out = absEnergy;
    return 0;
}

int orc_threebody_compute(const orc_threebody_force* f, const double* positions, const double* box, double lambda,
                          double* forces, double* energy) {
    return run_threebody(f, positions, box, lambda, forces, energy, NULL, NULL);
}

int orc_threebody_compute_scale(const orc_threebody_force* f, const double* positions, const double* box, double lambda,
                                double* forces, double* energy, double* abs_energy) {
    return run_threebody(f, positions, box, lambda, forces, energy, NULL, abs_energy);
}

long long orc_threebody_count(const orc_threebody_force* f, const double* positions, const double* box) {
    long long count = 0;
    if (run_threebody(f, positions, box, 1.0, NULL, NULL, &count, NULL) != 0)
        return -1;
    return count;
}
