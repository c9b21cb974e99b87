/*
 * FP64 CPU oracle of the InteraxFF pair and three-body forces.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference plugin's double precision Reference
 * platform, used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.  The HIP
 * product path never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the reference repository root).
 */
#ifndef INTERAX_ORACLE_H
#define INTERAX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Where the Slater overlap uses its Taylor expansion (InteraxPairTerms.h:132-160). */
typedef struct {
    int order;
    int about_smaller_width;
    double max_delta, max_t, max_tx;
} orc_taylor_scheme;

void orc_scheme_accurate(int double_precision, orc_taylor_scheme* scheme);
void orc_scheme_order2(orc_taylor_scheme* scheme);

/* Overlap integral S(r; a, b) and dS/dr (InteraxPairTerms.h:167-216). */
void orc_slater_overlap(double r, double a, double b, const orc_taylor_scheme* scheme, double* overlap, double* deriv);
/* Tang-Toennies f_6, f_8, f_10 and their x derivatives (InteraxPairTerms.h:223-253). */
void orc_tang_toennies(double x, double damping[3], double deriv[3]);
/* Pair energies and dE/dr (InteraxPairTerms.h:267-337).  Parameter rows as in orc_pair_force.params. */
void orc_dispersion_pair(double r, const double* pi, const double* pj, double damping_scale, double* energy, double* dedr);
void orc_exchange_pair(double r, const double* pi, const double* pj, const orc_taylor_scheme* s, double* energy, double* dedr);
void orc_short_range_pair(double r, const double* pi, const double* pj, const orc_taylor_scheme* s, double* energy, double* dedr);

enum { ORC_DISPERSION = 0, ORC_EXCHANGE = 1, ORC_SHORT_RANGE = 2 };
enum { ORC_NO_CUTOFF = 0, ORC_CUTOFF_NON_PERIODIC = 1, ORC_CUTOFF_PERIODIC = 2 };

/*
 * One pair force, as its Force object holds it (InteraxPairForce.h:70-239 and the subclasses).
 *   params: per particle, row-major: dispersion (m1, m2, m3, alpha, width), exchange (amplitude, width),
 *           short range (k, l, width).
 */
typedef struct {
    int kind;
    int num_particles;
    const double* params;
    const unsigned char* alchemical;
    int method;
    double cutoff;
    int use_switch;
    double switching_distance;
    double softcore_radius;
    int num_exclusions;
    const int* exclusions;      /* 2*num_exclusions particle indices */
    double damping_scale;       /* dispersion */
    int use_long_range_correction;
    int overlap_method;         /* 0 Exact, 2 InteraxOrder2 */
    int double_precision;       /* which Exact switch: the double or the single precision one */
} orc_pair_force;

/*
 * Energy of a pair force, forces ADDED to forces[3*n] (ReferenceInteraxPairIxn.h:53-99 with the kernel callbacks of
 * ReferenceInteraxKernels.cpp:79-183).  box: 3 row vectors (9 doubles), used by CutoffPeriodic.  The dispersion long
 * range correction is included, as in ReferenceInteraxKernels.cpp:93-94.  Returns 0 on success, -1 on bad input.
 */
int orc_pair_compute(const orc_pair_force* force, const double* positions, const double* box, double lambda,
                     double* forces, double* energy);
/* The same, also returning the sum of |pair energies| before switching, the scale of rounding errors that
   TestInteraxPairForce.h:90-127 compares energies against. */
int orc_pair_compute_scale(const orc_pair_force* force, const double* positions, const double* box, double lambda,
                           double* forces, double* energy, double* abs_energy);
/* Number of interacting pairs (not excluded, gate nonzero, r < cutoff) the same call would evaluate. */
long long orc_pair_count(const orc_pair_force* force, const double* positions, const double* box);
/* Long range correction coefficients (InteraxDispersionForceImpl.cpp:86-149). */
void orc_long_range_correction(const orc_pair_force* force, double* coefficient, double* alchemical_coefficient);
/* Gauss-Legendre rule on (0, 1) (InteraxDispersionForceImpl.cpp:49-73). */
void orc_gauss_legendre(int n, double* nodes, double* weights);

/*
 * Three-body force (InteraxThreeBodyForce.h:75-258).  Sites are centroids of particles listed in
 * site_particles[site_offsets[i] .. site_offsets[i+1]).
 */
typedef struct {
    int num_sites;
    const int* site_offsets;
    const int* site_particles;
    const double* alpha;
    const double* frequency;
    const double* width;
    const int* molecule;
    const unsigned char* alchemical;
    double cutoff, scale_factor, damping_scale, centroid_damping_radius;
    int damping_method;         /* 0 NoDamping, 1 TangToennies */
    int periodic;
} orc_threebody_force;

/* Tang-Toennies f_3 and centroid damping (InteraxThreeBodyTerms.h:51-77). */
void orc_tang_toennies3(double x, double* damping, double* deriv);
/* One ATM triple; forces ADDED (InteraxThreeBodyTerms.h:97-146).  rates NULL for no Tang-Toennies damping. */
double orc_triple(const double a[3], const double b[3], const double c[3], double c9, const double* rates,
                  double fa[3], double fb[3], double fc[3], double damping_radius);
double orc_c9(double a1, double w1, double a2, double w2, double a3, double w3);
/* ReferenceInteraxThreeBodyIxn.cpp:84-135.  Returns 0 on success. */
int orc_threebody_compute(const orc_threebody_force* force, const double* positions, const double* box, double lambda,
                          double* forces, double* energy);
/* The same, also returning the sum of |triple energies| (TestInteraxThreeBodyForce.h:150-165 compares against it). */
int orc_threebody_compute_scale(const orc_threebody_force* force, const double* positions, const double* box, double lambda,
                                double* forces, double* energy, double* abs_energy);
/* Number of interacting triples. */
long long orc_threebody_count(const orc_threebody_force* force, const double* positions, const double* box);

#ifdef __cplusplus
}
#endif
#endif
