/*
 * interax_hip: the C ABI of the MI355X (gfx950) HIP platform for the InteraxFF pair and three-body forces.
 *
 * This is the boundary a host binding calls.  It replaces, for one platform, the four kernel interfaces of the
 * reference plugin (openmmapi/include/InteraxKernels.h:51-189):
 *
 *   CalcInteraxDispersionForceKernel  (InteraxKernels.h:51-81)   -> ixh_pair_* with IXH_TERM_DISPERSION
 *   CalcInteraxExchangeForceKernel    (InteraxKernels.h:87-117)  -> ixh_pair_* with IXH_TERM_EXCHANGE
 *   CalcInteraxShortRangeForceKernel  (InteraxKernels.h:123-153) -> ixh_pair_* with IXH_TERM_SHORT_RANGE
 *   CalcInteraxThreeBodyForceKernel   (InteraxKernels.h:159-189) -> ixh_threebody_*
 *
 * Each interface's initialize() maps to *_create(), execute() to *_execute() and copyParametersToContext() to
 * *_set_parameters().  Pair forces that share the nonbonded method, cutoff, switching function, soft-core radius,
 * exclusions and alchemical flags are fused into one handle and evaluated by one kernel; the `terms` argument of
 * ixh_pair_execute() selects which of them an evaluation includes (the force groups of the reference's
 * getState(groups) call).
 *
 * Conventions.
 *   - Every function returns IXH_OK (0) or an error code; ixh_last_error() gives the message of the last error on
 *     the calling thread.  No C++ exception crosses this boundary.
 *   - Host arrays passed in are copied; the caller keeps ownership.  The library owns the device memory of its
 *     handles.  The device buffers of ixh_buffers are borrowed for the duration of one call.
 *   - A handle is not re-entrant: calls on one handle must not overlap.  Distinct handles may run concurrently on
 *     distinct streams.
 *   - Units are those of OpenMM: nm, kJ/mol, kJ/mol/nm.
 *   - Forces are accumulated in 64-bit fixed point with 2^32 units per kJ/mol/nm, as OpenMM's GPU platforms do,
 *     which makes them independent of the order of accumulation.  Energies are accumulated in FP64.
 */
#ifndef INTERAX_HIP_H
#define INTERAX_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IXH_VERSION_MAJOR 0
#define IXH_VERSION_MINOR 1

/* Status codes. */
#define IXH_OK 0
#define IXH_ERROR_INVALID_ARGUMENT 1   /* a setting or parameter the reference rejects (same messages) */
#define IXH_ERROR_HIP 2                /* a HIP runtime error */
#define IXH_ERROR_STATE 3              /* e.g. "The number of particles has changed" */
#define IXH_ERROR_UNSUPPORTED 4
#define IXH_ERROR_CAPACITY 5           /* a request that does not fit the device memory budget */

/* Precision of the evaluation: OpenMM's "Precision" platform property (single, mixed, double). */
#define IXH_SINGLE 0    /* FP32 positions and arithmetic */
#define IXH_MIXED 1     /* FP64 positions (float4 posq plus float4 correction), FP32 arithmetic */
#define IXH_DOUBLE 2    /* FP64 positions and arithmetic */

/* Nonbonded methods (InteraxPairForce.h:76-93). */
#define IXH_NO_CUTOFF 0
#define IXH_CUTOFF_NON_PERIODIC 1
#define IXH_CUTOFF_PERIODIC 2

/* Pair terms, one per pair force class. */
#define IXH_TERM_DISPERSION 1
#define IXH_TERM_EXCHANGE 2
#define IXH_TERM_SHORT_RANGE 4

/* Overlap methods (InteraxExchangeForce.h OverlapMethod). */
#define IXH_OVERLAP_EXACT 0
#define IXH_OVERLAP_INTERAX_ORDER2 2

/* Three-body damping methods (InteraxThreeBodyForce.h DampingMethod). */
#define IXH_DAMPING_NONE 0
#define IXH_DAMPING_TANG_TOENNIES 1

/* execute() flags */
#define IXH_INCLUDE_FORCES 1
#define IXH_INCLUDE_ENERGY 2
/* Pair forces with a cutoff, and three-body forces of plain handles: check the neighbour list against the positions on
   the device without waiting for the answer; the evaluation proceeds with the current list.  ixh_pair_list_was_valid()
   (ixh_threebody_list_was_valid()) then tells whether it was still valid (if not, the results must be discarded and the
   evaluation repeated without this flag, which rebuilds the list). */
#define IXH_DEFER_LIST_CHECK 4

/* Message of the last failed call on this thread ("" if none). */
const char* ixh_last_error(void);
/* Library version. */
int ixh_version(int* major, int* minor);
/* Number of visible HIP devices. */
int ixh_device_count(int* count);

/*
 * The per-Context device buffers of a platform: what OpenMM's platform data owns for a Context and hands to a
 * kernel's execute() (the posq array, the fixed-point force buffer, the energy accumulator, the box, the context
 * parameter and the atom index array).  All pointers are device pointers.
 *
 * Particle order.  posq and force are in the platform's current order, which may differ from the System's: OpenMM's
 * GPU platforms reorder atoms for locality, but only ever exchange whole molecules that are identical in every
 * parameter of every force (the ComputeForceInfo contract, CommonInteraxKernels.cpp:111-137, 386-412).  Per-particle
 * parameters, exclusions and three-body sites given in System order therefore stay valid in the current order; the
 * one thing that changes is the original index of the particle in each slot, which the InteraxOrder2 overlap rule
 * uses (the width of the particle with the lower original index, CommonInteraxKernels.cpp:213-216,
 * interaxSlaterOverlap.cc:14-17).  atom_index carries it: atom_index[i] is the System index of the particle held in
 * slot i (OpenMM's ComputeContext::getAtomIndexArray); NULL means the identity.  It is read on every call.
 */
typedef struct {
    int num_particles;
    int precision;                 /* IXH_SINGLE / IXH_MIXED: posq is float4; IXH_DOUBLE: posq is double4 */
    const void* posq;              /* x, y, z, (unused) per particle, in the platform's current order */
    const void* posq_correction;   /* IXH_MIXED: float4 low-order parts, position = posq + correction; else NULL */
    long long* force;              /* component c of slot i at force[c*force_stride + i], 2^32 per kJ/mol/nm */
    int force_stride;              /* >= num_particles */
    double* energy;                /* the energy of the call is ADDED to energy[0] */
    double box[9];                 /* periodic box vectors a, b, c (rows), in OpenMM's reduced form */
    double lambda;                 /* current value of the alchemical context parameter */
    void* stream;                  /* hipStream_t, or NULL for the default stream */
    const int* atom_index;         /* System index of the particle in each slot, or NULL (identity) */
} ixh_buffers;

/*
 * Device state: a stand-in for a platform's per-Context buffers, for hosts that do not bring their own (tests,
 * benchmarks, the server).  It allocates the buffers of ixh_buffers and converts positions and forces.
 */
typedef struct ixh_state ixh_state;
int ixh_state_create(int device, int num_particles, int precision, ixh_state** out);
int ixh_state_destroy(ixh_state* state);
/* Upload positions (3 doubles per particle). */
int ixh_state_set_positions(ixh_state* state, const double* positions);
/* Set the periodic box vectors (9 doubles, rows a, b, c). */
int ixh_state_set_box(ixh_state* state, const double* box);
/* Set the value of the alchemical context parameter used by the next calls. */
int ixh_state_set_lambda(ixh_state* state, double lambda);
/* Run the state's operations, and hand out in ixh_buffers, the given stream (hipStream_t; NULL: the state's own) in
   place of the state's own, as a platform runs all kernels of a Context on one stream (OpenMM's
   CudaContext::getCurrentStream).  Waits for the work queued on the previous stream.  The stream must outlive its
   use by the state. */
int ixh_state_set_stream(ixh_state* state, void* stream);
/* Zero the force buffer and the energy accumulator (asynchronous, on the state's stream). */
int ixh_state_clear(ixh_state* state);
/* Wait for the state's stream and download the forces (3 doubles per particle) / the energy. */
int ixh_state_get_forces(ixh_state* state, double* forces);
int ixh_state_get_energy(ixh_state* state, double* energy);
/* Wait for all work queued on the state's stream. */
int ixh_state_synchronize(ixh_state* state);
/* The buffers of the state, to pass to the execute() functions. */
int ixh_state_get_buffers(ixh_state* state, ixh_buffers* out);

/*
 * A fused set of pair forces that share their geometry settings (InteraxPairForce.h:70-239).  The per-particle
 * parameter arrays of the terms not in `terms` may be NULL.
 */
typedef struct {
    int num_particles;
    int terms;                      /* IXH_TERM_* bitmask of the forces in this handle */
    int method;                     /* IXH_NO_CUTOFF, IXH_CUTOFF_NON_PERIODIC, IXH_CUTOFF_PERIODIC */
    double cutoff;
    int use_switching_function;
    double switching_distance;
    double softcore_radius;
    const unsigned char* alchemical;   /* num_particles flags, or NULL for none */
    int num_exclusions;
    const int* exclusions;             /* 2*num_exclusions particle indices */
    int precision;                     /* IXH_SINGLE, IXH_MIXED or IXH_DOUBLE */
    /* InteraxDispersionForce: per particle (m1, m2, m3, alpha, width) */
    const double* dispersion;
    double damping_scale;
    int use_long_range_correction;
    /* InteraxExchangeForce: per particle (amplitude, width) */
    const double* exchange;
    int exchange_overlap_method;
    /* InteraxShortRangeForce: per particle (k, l, width) */
    const double* short_range;
    int short_range_overlap_method;
    /* Periodic box used to validate the cutoff (the System's default box); may be NULL if not periodic. */
    const double* default_box;
} ixh_pair_desc;

typedef struct ixh_pair ixh_pair;
/* CalcInterax*ForceKernel::initialize (ReferenceInteraxKernels.cpp:44-50, 104-112, 147-155). */
int ixh_pair_create(int device, const ixh_pair_desc* desc, ixh_pair** out);
int ixh_pair_destroy(ixh_pair* pair);
/*
 * CalcInterax*ForceKernel::copyParametersToContext (ReferenceInteraxKernels.cpp:98-112, 142-148, 185-191): new
 * per-particle parameters and damping scale.  The number of particles and the overlap methods must not change.
 * The long range correction is recomputed.  Synchronous: the host arrays may be reused on return.
 */
int ixh_pair_set_parameters(ixh_pair* pair, const ixh_pair_desc* desc);
/*
 * The same update from per-particle arrays in device memory (no counterpart in the reference, whose
 * copyParametersToContext reads the Force's host-side parameters, ReferenceInteraxKernels.cpp:98-112, 142-148,
 * 185-191): for callers whose parameters live on the GPU (fitting loops, alchemical schedules).  The arrays are in the
 * System's particle order and the descriptor's layout: dispersion m1, m2, m3, alpha, width (5 doubles per particle),
 * exchange amplitude, width (2), short range K, L, width (3); NULL for a term the handle lacks.  damping_scale is
 * ignored without dispersion.  The call is asynchronous on `stream`: it copies the arrays there (keep them valid until
 * that work has run) and rederives, on the device, the evaluation's arrays, their copies in the neighbour list's order
 * (the list is kept) and the long range correction; it waits for nothing.  The next evaluation, on any stream, is
 * ordered after it.  Two cases make that evaluation wait for the update on the host first: type tables (SINGLE and
 * MIXED precision with few parameter classes: the update is checked against the classes, the table rows are formed
 * from the classes' parameters, and the classes are formed anew only if a particle no longer matches its class) and
 * one overlap shared by exchange and short range (the widths must still agree).  ixh_pair_long_range_* wait for the
 * update's correction.  ixh_pair_set_parameters runs the same device work, so both give the same results bit for
 * bit.  ixh_pair_execute_noalloc returns IXH_ERROR_STATE while an update that needs the host is pending.
 */
typedef struct {
    const double* dispersion;
    const double* exchange;
    const double* short_range;
    double damping_scale;
    void* stream;   /* hipStream_t */
} ixh_pair_device_parameters;
int ixh_pair_set_parameters_device(ixh_pair* pair, const ixh_pair_device_parameters* parameters);
/*
 * CalcInterax*ForceKernel::execute (ReferenceInteraxKernels.cpp:79-96, 127-140, 170-183): add the forces and
 * energy of the selected terms to the buffers.  The long range correction is not included: it is returned by
 * ixh_pair_long_range_energy(), as the GPU platforms return it from execute().  Asynchronous on buffers->stream.
 */
int ixh_pair_execute(ixh_pair* pair, const ixh_buffers* buffers, int terms, int flags);
/*
 * ixh_pair_execute without any allocation (no counterpart in the reference; for callers that cannot allocate: code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The same kernels and launch as
 * ixh_pair_execute, with the same results bit for bit; the per-workgroup energy and pair-count partials of the launch
 * go to a static device array of the library (65 536 partials, 768 KiB per GPU, held until the process ends) instead of
 * the handle's buffers, and the call creates no memory, event or stream.  Limits:
 *   - the handle's neighbour list must exist: the first evaluation of a handle is made by ixh_pair_execute
 *     (IXH_ERROR_STATE otherwise);
 *   - a cutoff handle whose list no longer holds for the positions returns IXH_ERROR_STATE without evaluating (without
 *     IXH_DEFER_LIST_CHECK; with it, ixh_pair_list_was_valid tells as for ixh_pair_execute); ixh_pair_execute then
 *     rebuilds the list;
 *   - lists that need more partials than the array holds (more than 262 144 tiles with a cutoff, 65 536 without)
 *     return IXH_ERROR_CAPACITY; segmented handles return IXH_ERROR_UNSUPPORTED.
 * Every call on one GPU shares that GPU's array (each launch writes all the partials it reads, so nothing needs
 * clearing): calls that overlap on one GPU, from other threads or on other streams, race.
 */
int ixh_pair_execute_noalloc(ixh_pair* pair, const ixh_buffers* buffers, int terms, int flags);
/* Dispersion long range correction: coefficients (InteraxDispersionForceImpl.cpp:86-149) and the energy
   (coefficient + lambda*alchemicalCoefficient)/V (ReferenceInteraxKernels.cpp:93-94). */
int ixh_pair_long_range_coefficients(ixh_pair* pair, double* coefficient, double* alchemical_coefficient);
int ixh_pair_long_range_energy(ixh_pair* pair, const double* box, double lambda, double* energy);
/*
 * Energies of the selected terms at count (1 to 16) values of lambda_sterics, at the positions of the buffers
 * (lambda-window fusion; no counterpart in the reference, which evaluates one lambda per Context: a free-energy
 * analysis needs the energy of one configuration at every window).  Only pairs with exactly one alchemical particle
 * depend on lambda (ReferenceInteraxPairIxn.h:100-125): one evaluation at buffers->lambda (into the handle's own
 * scratch buffers: the buffers' forces and energy are not touched), then those pairs at every lambda.  The long range
 * correction is not included (ixh_pair_long_range_energy).  Waits for the result.
 */
int ixh_pair_lambda_energies(ixh_pair* pair, const ixh_buffers* buffers, int terms, const double* lambdas, int count, double* energies);
/*
 * Batches of independent systems (no counterpart in the reference; used by the batched and served forms): the
 * particles of num_segments systems are concatenated, segment s owning particles [particle_offsets[s],
 * particle_offsets[s+1]); pairs never cross segments.  The descriptor describes the concatenated system (exclusions in
 * global indices); the NoCutoff method only.  ixh_pair_execute_segments also ADDS the energy of every segment to
 * segment_energies, a device array of num_segments doubles (the per-segment sums have a fixed order, so a segment's
 * energy and forces do not depend on the other segments of the batch).
 */
int ixh_pair_create_segmented(int device, const ixh_pair_desc* desc, int num_segments, const int* particle_offsets, ixh_pair** pair);
int ixh_pair_execute_segments(ixh_pair* pair, const ixh_buffers* buffers, int terms, int flags, double* segment_energies);
/* The same for the listed segments only (host array of num_selected indices); the others are not evaluated and their
   forces and energies are left as they are. */
int ixh_pair_execute_subset(ixh_pair* pair, const ixh_buffers* buffers, int terms, int flags, double* segment_energies,
                            const int* segments, int num_selected);
/* Whether the neighbour list used by the last execute() with IXH_DEFER_LIST_CHECK was valid for its positions
   (waits for the check); 1 for any other execute(). */
int ixh_pair_list_was_valid(ixh_pair* pair, int* valid);
/* Number of interacting pairs of the last execute() (waits for it). */
int ixh_pair_last_pair_count(ixh_pair* pair, long long* count);
/* Launch shape of the cutoff tile kernel (no counterpart in the reference): each wave takes a work item of up to
   tiles_per_item consecutive tiles of one 64-particle block (its i side staged and its forces flushed once per item).
   Takes effect at the next evaluation, which rebuilds the neighbour list; results do not depend on it beyond the
   order of the FP64 energy sum. */
int ixh_pair_set_tiles_per_item(ixh_pair* pair, int tiles_per_item);
/* Tuner of that shape (no counterpart in the reference): times 1, 2, 3, 4 and 6 tiles per work item on the positions
   of the buffers (evaluating into scratch buffers of the handle; the buffers' forces and energy are not touched) and
   keeps the fastest, which is returned in tiles_per_item (may be NULL).  A handle with a cutoff only; waits for the
   device. */
int ixh_pair_autotune(ixh_pair* pair, const ixh_buffers* buffers, int terms, int* tiles_per_item);
/* How the current neighbour list takes periodic images (no counterpart in the reference): of its `tiles` tiles,
   `per_block_tiles` take one image shift per pair of 64-particle blocks (exact for every pair in range when the box is
   wider than twice the cutoff plus the extents of both blocks and the skin); the others take the minimum image of
   every pair.  0 per-block tiles when the method is not periodic.  Waits for the last execute(). */
int ixh_pair_image_mode(ixh_pair* pair, long long* per_block_tiles, long long* tiles);
/* Size of the current neighbour list (no counterpart in the reference): the number of 64 x 64 tiles and of candidate
   pairs in them (within the cutoff plus the list skin, exclusions removed).  Waits for the last execute(). */
int ixh_pair_list_statistics(ixh_pair* pair, long long* tiles, long long* candidates);

/*
 * Measurement (no counterpart in the reference): with profiling enabled, every launch of the pair kernel (and of
 * the triple kernel) is bracketed by HIP events on its stream.  *_kernel_time() waits for the recorded launches and
 * returns their summed duration in milliseconds and their number since the last reset.
 */
int ixh_set_profiling(int enable);
int ixh_pair_kernel_time(ixh_pair* pair, double* milliseconds, long long* launches);
int ixh_pair_reset_kernel_time(ixh_pair* pair);

/* InteraxThreeBodyForce (InteraxThreeBodyForce.h:75-258). */
typedef struct {
    int num_particles;                 /* of the System */
    int num_sites;
    const int* site_offsets;           /* num_sites+1 offsets into site_particles */
    const int* site_particles;
    const double* alpha;
    const double* frequency;
    const double* width;
    const int* molecule;               /* negative: a molecule of its own */
    const unsigned char* alchemical;   /* or NULL */
    double cutoff;
    double scale_factor;
    int damping_method;                /* IXH_DAMPING_NONE or IXH_DAMPING_TANG_TOENNIES */
    double damping_scale;
    double centroid_damping_radius;
    int periodic;
    int precision;
} ixh_threebody_desc;

typedef struct ixh_threebody ixh_threebody;
/* CalcInteraxThreeBodyForceKernel::initialize, with the checks of InteraxThreeBodyForceImpl.cpp:42-62. */
int ixh_threebody_create(int device, const ixh_threebody_desc* desc, ixh_threebody** out);
int ixh_threebody_destroy(ixh_threebody* threebody);
/* copyParametersToContext (ReferenceInteraxKernels.cpp:215-217): alpha, frequency, width, scale factor, damping
   scale and a nonzero centroid damping radius may change. */
int ixh_threebody_set_parameters(ixh_threebody* threebody, const ixh_threebody_desc* desc);
/*
 * The same update from per-site arrays in device memory (no counterpart in the reference, whose copyParametersToContext
 * reads the Force's host-side parameters, ReferenceInteraxKernels.cpp:215-217): alpha, frequency and width, FP64, one
 * per site in site order, plus the scalars.  Asynchronous on `stream` (keep the arrays valid until that work has run);
 * the neighbour lists are kept.  The next evaluation, on any stream, is ordered after it; with Tang-Toennies damping
 * that evaluation first waits for the update's check of the widths on the device and fails with
 * IXH_ERROR_INVALID_ARGUMENT for a width that is not positive.  ixh_threebody_set_parameters runs the same device work.
 */
typedef struct {
    const double* alpha;
    const double* frequency;
    const double* width;
    double scale_factor;
    double damping_scale;
    double centroid_damping_radius;   /* may change, but not between zero and nonzero */
    void* stream;                     /* hipStream_t */
} ixh_threebody_device_parameters;
int ixh_threebody_set_parameters_device(ixh_threebody* threebody, const ixh_threebody_device_parameters* parameters);
/* execute (ReferenceInteraxKernels.cpp:210-213). */
int ixh_threebody_execute(ixh_threebody* threebody, const ixh_buffers* buffers, int flags);
/* Batches of independent nonperiodic systems (no counterpart in the reference): segment s owns sites
   [site_offsets[s], site_offsets[s+1]); triples never cross segments; the per-segment energies are ADDED to the device
   array segment_energies.  An evaluation of a segmented handle needs no host synchronization. */
int ixh_threebody_create_segmented(int device, const ixh_threebody_desc* desc, int num_segments, const int* site_offsets,
                                   ixh_threebody** threebody);
int ixh_threebody_execute_segments(ixh_threebody* threebody, const ixh_buffers* buffers, int flags, double* segment_energies);
int ixh_threebody_execute_subset(ixh_threebody* threebody, const ixh_buffers* buffers, int flags, double* segment_energies,
                                 const int* segments, int num_selected);
/* Plain handles keep neighbour lists within cutoff + 0.1 nm and rebuild them when a site has moved more than half of
   that skin (interaxThreeBody.cc:11-30).  execute() checks the lists on the device; with IXH_DEFER_LIST_CHECK it does
   not wait for the answer, and this call then tells whether the lists of the last execute() held (if not, its results
   must be discarded and the evaluation repeated without the flag, which rebuilds them); 1 for any other execute(). */
int ixh_threebody_list_was_valid(ixh_threebody* threebody, int* valid);
/* Number of interacting triples of the last execute() (waits for it). */
int ixh_threebody_last_triple_count(ixh_threebody* threebody, long long* count);
int ixh_threebody_kernel_time(ixh_threebody* threebody, double* milliseconds, long long* launches);
int ixh_threebody_reset_kernel_time(ixh_threebody* threebody);

/*
 * Batched evaluation of many independent systems (no counterpart in the reference: its kernels evaluate one System per
 * Context).  The systems are concatenated: system s owns particles [particle_offsets[s], particle_offsets[s+1]) and
 * three-body sites [site_offsets[s], site_offsets[s+1]); `pair` describes the concatenated pair forces (NoCutoff,
 * global particle indices) and `threebody` the concatenated nonperiodic three-body force; either may be NULL.  The
 * batch owns its device state.  ixh_batch_execute is asynchronous; from the second call on it replays a HIP graph of
 * the whole evaluation.
 */
typedef struct ixh_batch ixh_batch;
typedef struct {
    int num_systems;
    const int* particle_offsets;
    const ixh_pair_desc* pair;
    const int* site_offsets;
    const ixh_threebody_desc* threebody;
    int precision;
} ixh_batch_desc;
int ixh_batch_create(int device, const ixh_batch_desc* desc, ixh_batch** batch);
int ixh_batch_destroy(ixh_batch* batch);
int ixh_batch_set_positions(ixh_batch* batch, const double* positions);   /* host, 3 x total particles, nm */
int ixh_batch_get_buffers(ixh_batch* batch, ixh_buffers* buffers);      /* the batch's device state */
int ixh_batch_execute(ixh_batch* batch, int flags);
int ixh_batch_get_energies(ixh_batch* batch, double* energies);          /* waits; one per system, kJ/mol */
int ixh_batch_get_forces(ixh_batch* batch, double* forces);              /* waits; 3 x total particles */
int ixh_batch_set_graph(ixh_batch* batch, int enable);                   /* 0: launch the kernels eagerly */
/*
 * Launch shape of a batch's evaluation: the workgroups of the persistent pair-tile launch, and the triple kernel's
 * wavefronts per work item (1, 2, 4 or 8) and neighbour pairs per work item (a multiple of 64 in [1024, 2^22]; a
 * work item is a site and consecutive shares of its neighbour pairs, staged once in LDS for its wavefronts).  Every
 * shape gives the same per-system energies and forces bit for bit: the pair tiles and the shares (fixed by a site's
 * list capacity, one wavefront each) add their forces in fixed point and their energies per tile or share, summed per
 * system in a fixed order.  ixh_batch_autotune times the pair launch's workgroup counts, then the
 * triple kernel's shapes with the chosen pair launch, on the batch's current positions, and keeps the fastest
 * (the set and the choice are in chosen, which may be NULL).  Setting a shape waits for the device.
 */
typedef struct {
    int pair_workgroups;
    int triple_waves;
    int triple_pairs_per_item;
} ixh_batch_shape;
int ixh_batch_get_shape(ixh_batch* batch, ixh_batch_shape* shape);
int ixh_batch_set_shape(ixh_batch* batch, const ixh_batch_shape* shape);
int ixh_batch_autotune(ixh_batch* batch, int flags, ixh_batch_shape* chosen);
/* With profiling enabled (ixh_set_profiling), the summed durations of the batch's pair-tile and triple kernels since
   the last reset (eager evaluations only: profiling evaluates eagerly), and their launch counts. */
int ixh_batch_kernel_time(ixh_batch* batch, double* pair_ms, long long* pair_launches, double* triple_ms, long long* triple_launches);
int ixh_batch_reset_kernel_time(ixh_batch* batch);

/*
 * Serving (no counterpart in the reference): a request loop over device-resident systems, with its own thread.
 * Register systems, start, then submit requests and wait for their tickets; systems may also be registered while the
 * server runs.  Cluster systems (NoCutoff pair forces and an optional nonperiodic three-body force, all with the same
 * settings) form a catalog that is evaluated in batches: whenever one of two execution lanes is free, the pending
 * cluster requests are admitted in arrival order up to max_batch requests and the memory budget; a system with several
 * requests in a batch is evaluated in rounds of one request per system.  A cluster registered while the server runs is
 * appended to a new catalog, built by the registering call while the lanes serve the old one and taken up by each lane
 * when it is next idle; its requests wait until then.  Periodic systems (cutoff methods, e.g. alchemical boxes) have
 * execution slots of their own (a handle and a state each, one request per slot, all on one stream): a request at one
 * lambda (energy and forces), or, through ixh_server_submit_windows, the energies of one configuration at up to 16
 * values of lambda by the fused lambda-window evaluation (ixh_pair_lambda_energies; long range correction included).
 * The optional positions of ixh_server_add_periodic (3 per particle) are evaluated once at registration, which builds
 * the slots' neighbour lists and sizes their buffers.  A request's result does not depend on the other requests of
 * its batch.
 * Memory: memory_budget bounds the memory of the requests in flight.  A cluster request holds its share of its lane's
 * staging (pinned host memory that the scatter and gather kernels read and write directly: 56 bytes per particle and
 * 24 per request); a periodic request holds its slot, whose device memory is measured when the slot is created (the
 * library's own allocations).  A request is admitted when it fits beside
 * those in flight, or when nothing is in flight.  ixh_server_memory reports the resident device memory of the catalogs
 * and slots, the bytes in flight and their peak.
 * ixh_server_replay submits n requests at the given arrival times (seconds from the call; window_counts, optional,
 * gives each request's number of lambda values, taken in order from window_lambdas) and returns each one's latency
 * (scheduled arrival to completion) and energy (a window request's first), and the time to the last completion.
 */
typedef struct ixh_server ixh_server;
int ixh_server_create(int device, int precision, int max_batch, long long memory_budget, ixh_server** server);
int ixh_server_add_cluster(ixh_server* server, const ixh_pair_desc* pair, const ixh_threebody_desc* threebody, int* entry);
int ixh_server_add_periodic(ixh_server* server, const ixh_pair_desc* pair, const double* box, const double* positions,
                            int* entry);
int ixh_server_start(ixh_server* server);
int ixh_server_submit(ixh_server* server, int entry, const double* positions, double lambda, long long* ticket);
int ixh_server_submit_windows(ixh_server* server, int entry, const double* positions, const double* lambdas, int count,
                              long long* ticket);
int ixh_server_wait(ixh_server* server, long long ticket, double* energy, double* forces, double* latency_ms);
int ixh_server_wait_windows(ixh_server* server, long long ticket, double* energies, double* latency_ms);
int ixh_server_statistics(ixh_server* server, long long* batches, long long* batched_requests);
int ixh_server_memory(ixh_server* server, long long* resident_bytes, long long* in_flight_bytes, long long* peak_in_flight_bytes);
int ixh_server_replay(ixh_server* server, int n, const double* arrival_s, const int* entries, const double* lambdas,
                      const long long* position_offsets, const double* positions, const int* window_counts,
                      const double* window_lambdas, double* latency_ms, double* energies, double* makespan_s);
int ixh_server_destroy(ixh_server* server);

#ifdef __cplusplus
}
#endif
#endif
