// A serving loop over device-resident systems (the served configuration: a Poisson stream of cluster evaluations
// and alchemical λ-window evaluations).
//
// No counterpart in the reference.  Systems are registered once; their parameters stay on the device.  Cluster
// systems form one catalog, concatenated into segmented pair and three-body handles, instantiated twice ("lanes",
// each with its own stream and state) so that two batches are on the GPU at once.  The server thread keeps a FIFO of
// pending requests; whenever a lane is free it admits, in arrival order, the pending cluster requests that fit the
// batch size and the memory budget, evaluates exactly their segments (in rounds when a system has several requests in
// the batch: a round holds at most one request per system) and signals the requests when the lane's completion event
// has passed.  The requests' coordinates and results move between pinned host memory and the device by the kernels that
// scatter and gather them (they read and write the host memory directly), not by the copy engines: under a stream of
// small requests, queued copies held up the submitting thread for milliseconds now and then.
// Periodic systems (alchemical boxes with a cutoff) have their own execution slots on a third stream, one request per
// slot: a request at one λ (energy and forces), or the energies of one configuration at up to 16 λ values through the
// fused λ-window evaluation.  A request's result does not depend on the other requests of its batch: segments are
// evaluated independently and their sums have a fixed order.
//
// Memory.  The budget bounds the memory of the requests in flight: a cluster request holds its share of its lane's
// pinned staging (packed coordinates, gathered results), a periodic request the device memory of its slot (its
// handle and state, as allocated).  A request is admitted when it fits beside those in flight, or alone.
//
// Systems may be registered while the server runs: a new periodic system gets its slots at once; a new cluster makes
// a new catalog (the old clusters keep their places, the new one is appended), built by the registering thread while
// the lanes serve the old one, and taken up by each lane when it is next idle.
#include "ixh_common.h"
#include "ixh_internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

using namespace ixh;

namespace {

typedef std::chrono::steady_clock Clock;
constexpr int MAX_SYSTEMS = 1 << 16;
constexpr int MAX_SLOTS = 1 << 16;

void check(int status) {
    if (status != IXH_OK)
        throw Error(status, ixh_last_error());
}

// Scatter the coordinates of a round, packed request after request, into the catalog's position buffers.
// items: (first particle, count, offset in the packed array) per request.
__global__ void k_scatter_positions(const float4* packed, const float4* packedCorrection, const int3* items, int numItems,
                                    float4* posq, float4* correction) {
    const int item = blockIdx.y;
    if (item >= numItems)
        return;
    const int3 it = items[item];
    for (int k = blockIdx.x*blockDim.x+threadIdx.x; k < it.y; k += gridDim.x*blockDim.x) {
        posq[it.x+k] = packed[it.z+k];
        if (correction != nullptr)
            correction[it.x+k] = packedCorrection[it.z+k];
    }
}

// Gather the forces (fixed point to kJ/mol/nm) of a round's requests and their segment energies into one array:
// per request 3 x count forces at 3 x offset, the energy of item k at energyOffset + k.
__global__ void k_gather_results(const long long* force, int stride, const int3* items, const int* segments, int numItems,
                                 const double* segmentEnergies, double* out, int energyOffset) {
    const int item = blockIdx.y;
    if (item >= numItems)
        return;
    const int3 it = items[item];
    for (int k = blockIdx.x*blockDim.x+threadIdx.x; k < it.y; k += gridDim.x*blockDim.x)
        for (int c = 0; c < 3; c++)
            out[3*(size_t) (it.z+k)+c] = (double) force[(size_t) c*stride+it.x+k]/FIXED_POINT_SCALE;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        out[energyOffset+item] = segmentEnergies[segments[item]];
}

// Coordinates of a periodic request from pinned host memory into its slot's state.
__global__ void k_copy_positions(const float4* posq, const float4* correction, int n, float4* posqOut, float4* correctionOut) {
    for (int k = blockIdx.x*blockDim.x+threadIdx.x; k < n; k += gridDim.x*blockDim.x) {
        posqOut[k] = posq[k];
        if (correctionOut != nullptr)
            correctionOut[k] = correction[k];
    }
}

__global__ void k_zero_doubles(double* p, int n) {
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n)
        p[i] = 0;
}

// Fixed-point forces of all particles and the energy of a periodic slot into one array.
__global__ void k_periodic_results(const long long* force, int stride, int n, const double* energy, double* out) {
    for (int k = blockIdx.x*blockDim.x+threadIdx.x; k < n; k += gridDim.x*blockDim.x)
        for (int c = 0; c < 3; c++)
            out[3*(size_t) k+c] = (double) force[(size_t) c*stride+k]/FIXED_POINT_SCALE;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        out[3*(size_t) n] = *energy;
}

struct Request {
    long long ticket = 0;
    int entry = 0;
    double lambda = 1;
    std::vector<double> lambdas;    // window requests: the energies at these values (no forces)
    std::vector<double> positions;
    long long bytes = 0;            // device memory held while in flight
    Clock::time_point arrival;
    Clock::time_point completion;
    bool done = false;
    int status = IXH_OK;
    std::string error;
    double energy = 0;
    std::vector<double> forces, energies;
};

// A registered system: a cluster of the catalog, or a periodic system with its own slots.
struct Entry {
    bool periodic = false;
    int numParticles = 0;
    // cluster: its segment and particle range in the catalog
    int segment = -1, firstParticle = 0;
    int numSites = 0;
    std::vector<double> dispersion, exchange, shortRange, alpha, frequency, width;
    std::vector<unsigned char> alchemical, siteAlchemical;
    std::vector<int> exclusions, siteOffsets, siteParticles, molecule;
    // periodic: the box, the slots of the system and the device memory of one slot
    double box[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<int> slots;
    long long slotBytes = 0;
};

// Memory of a cluster request in its lane's pinned staging: coordinates and corrections (2 float4), forces (3 doubles)
// per particle; its energy, item and segment.
long long cluster_request_bytes(int numParticles) {
    return 56LL*numParticles+(long long) (sizeof(double)+sizeof(int3)+sizeof(int));
}

// One execution slot of a periodic system: a handle, a state, staging and results; one request at a time.
struct PeriodicSlot {
    int entry = 0;
    ixh_pair* pair = nullptr;
    ixh_state* state = nullptr;
    ixh_buffers buffers;
    float4* hostPosq = nullptr;          // pinned staging, read and written by the kernels through devPosq etc.
    float4* hostCorrection = nullptr;
    double* hostResults = nullptr;       // forces (3n) and the energy
    float4* devPosq = nullptr;
    float4* devCorrection = nullptr;
    double* devResults = nullptr;
    hipEvent_t done = nullptr;
    std::shared_ptr<Request> request;
    ~PeriodicSlot() {
        if (pair != nullptr)
            ixh_pair_destroy(pair);
        if (state != nullptr)
            ixh_state_destroy(state);
        for (void* p : {(void*) hostPosq, (void*) hostCorrection, (void*) hostResults})
            if (p != nullptr)
                (void) hipHostFree(p);
        if (done != nullptr)
            (void) hipEventDestroy(done);
    }
};

// A lane's device copy of the cluster catalog: the first numClusters registered clusters.
struct Catalog {
    int numClusters = 0, numParticles = 0;
    ixh_pair* pair = nullptr;
    ixh_threebody* threebody = nullptr;
    ixh_state* state = nullptr;
    ixh_buffers buffers;
    DeviceArray<double> energies;   // one per segment
    long long bytes = 0;            // device memory of the handles and the state
    ~Catalog() {
        if (pair != nullptr)
            ixh_pair_destroy(pair);
        if (threebody != nullptr)
            ixh_threebody_destroy(threebody);
        if (state != nullptr)
            ixh_state_destroy(state);
    }
};

struct Lane {
    std::unique_ptr<Catalog> catalog, pending;   // pending: a larger catalog, taken up when the lane is idle
    hipStream_t stream = nullptr;
    // pinned staging (positions, corrections, items, segments, results), read and written by the kernels through the
    // device addresses dev*
    float4* hostPacked = nullptr;
    float4* hostPackedCorrection = nullptr;
    int3* hostItems = nullptr;
    int* hostSegments = nullptr;
    double* hostResults = nullptr;
    float4* devPacked = nullptr;
    float4* devPackedCorrection = nullptr;
    int3* devItems = nullptr;
    int* devSegments = nullptr;
    double* devResults = nullptr;
    size_t capacityParticles = 0, capacityItems = 0;
    hipEvent_t done = nullptr;
    std::vector<std::shared_ptr<Request>> batch;   // in round order
    int batchParticles = 0;
    bool busy = false;
};

}  // namespace

struct ixh_server {
    int device = 0;
    int precision = IXH_MIXED;
    int maxBatch = 64;
    long long memoryBudget = 0;
    // registered systems: entries[0 .. numEntries) are published (the vector never reallocates)
    std::vector<std::unique_ptr<Entry>> entries;
    std::atomic<int> numEntries{0};
    int numClusters = 0, catalogParticles = 0;   // registered clusters (under addMutex)
    // settings shared by the catalog's clusters (from the first registered cluster)
    ixh_pair_desc pairTemplate;
    ixh_threebody_desc threebodyTemplate;
    bool haveThreebody = false;
    bool haveClusterTemplate = false;
    Lane lanes[2];
    hipStream_t periodicStream = nullptr;
    std::vector<std::unique_ptr<PeriodicSlot>> periodicSlots;   // [0 .. numSlots) published
    std::atomic<int> numSlots{0};
    int slotsPerPeriodic = 4;
    std::thread worker;
    std::mutex mutex;      // requests, tickets, pending catalogs, statistics
    std::mutex addMutex;   // registrations (one at a time)
    std::condition_variable submitted, completed;
    std::deque<std::shared_ptr<Request>> incoming;
    std::map<long long, std::shared_ptr<Request>> tickets;
    long long nextTicket = 1;
    bool started = false, stopping = false;
    long long batches = 0, batchedRequests = 0, rounds = 0;
    long long residentBytes = 0;   // catalogs and slots
    long long inFlightBytes = 0, peakInFlightBytes = 0;
    std::string fatal;
};

namespace {

void release_lane_staging(Lane& lane) {
    for (void* p : {(void*) lane.hostPacked, (void*) lane.hostPackedCorrection, (void*) lane.hostItems, (void*) lane.hostSegments,
                    (void*) lane.hostResults})
        if (p != nullptr)
            (void) hipHostFree(p);
    lane.hostPacked = lane.hostPackedCorrection = nullptr;
    lane.hostItems = nullptr;
    lane.hostSegments = nullptr;
    lane.hostResults = nullptr;
}

// Pinned host memory that kernels read and write directly, and its device address.
template <class T>
void mapped_alloc(T*& host, T*& device, size_t count) {
    IXH_HIP_CHECK(hipHostMalloc((void**) &host, sizeof(T)*std::max<size_t>(count, 1), hipHostMallocMapped | hipHostMallocCoherent));
    void* d = nullptr;
    IXH_HIP_CHECK(hipHostGetDevicePointer(&d, host, 0));
    device = (T*) d;
}

void ensure_capacity(Lane& lane, size_t particles, size_t items) {
    if (particles > lane.capacityParticles || items > lane.capacityItems) {
        release_lane_staging(lane);
        lane.capacityParticles = std::max(lane.capacityParticles, std::max<size_t>(particles, 4096));
        lane.capacityItems = std::max(lane.capacityItems, std::max<size_t>(items, 256));
        mapped_alloc(lane.hostPacked, lane.devPacked, lane.capacityParticles);
        mapped_alloc(lane.hostPackedCorrection, lane.devPackedCorrection, lane.capacityParticles);
        mapped_alloc(lane.hostItems, lane.devItems, lane.capacityItems);
        mapped_alloc(lane.hostSegments, lane.devSegments, lane.capacityItems);
        mapped_alloc(lane.hostResults, lane.devResults, 3*lane.capacityParticles+lane.capacityItems);
    }
}

void split(double x, float& hi, float& lo) {
    hi = (float) x;
    lo = (float) (x-(double) hi);
}

// A catalog of the clusters among entries[0 .. count): concatenated in registration order into segmented handles and
// a state that runs on `stream`.
std::unique_ptr<Catalog> build_catalog(ixh_server* s, int count, hipStream_t stream) {
    std::unique_ptr<Catalog> c(new Catalog());
    std::vector<const Entry*> clusters;
    for (int k = 0; k < count; k++)
        if (!s->entries[k]->periodic)
            clusters.push_back(s->entries[k].get());
    if (clusters.empty())
        return c;
    std::vector<int> particleOffsets(clusters.size()+1, 0), siteOffsets(clusters.size()+1, 0);
    std::vector<double> dispersion, exchange, shortRange, alpha, frequency, width;
    std::vector<unsigned char> alchemical, siteAlchemical;
    std::vector<int> exclusions, siteParticleOffsets(1, 0), siteParticles, molecule;
    for (size_t k = 0; k < clusters.size(); k++) {
        const Entry& e = *clusters[k];
        particleOffsets[k+1] = particleOffsets[k]+e.numParticles;
        dispersion.insert(dispersion.end(), e.dispersion.begin(), e.dispersion.end());
        exchange.insert(exchange.end(), e.exchange.begin(), e.exchange.end());
        shortRange.insert(shortRange.end(), e.shortRange.begin(), e.shortRange.end());
        alchemical.insert(alchemical.end(), e.alchemical.begin(), e.alchemical.end());
        for (int x : e.exclusions)
            exclusions.push_back(x+particleOffsets[k]);
        siteOffsets[k+1] = siteOffsets[k]+e.numSites;
        for (int i = 0; i < e.numSites; i++) {
            for (int q = e.siteOffsets[i]; q < e.siteOffsets[i+1]; q++)
                siteParticles.push_back(e.siteParticles[q]+particleOffsets[k]);
            siteParticleOffsets.push_back((int) siteParticles.size());
            // negative molecule indices become -1 - (catalog site index)
            molecule.push_back(e.molecule[i] < 0 ? -1-(siteOffsets[k]+i) : e.molecule[i]);
        }
        alpha.insert(alpha.end(), e.alpha.begin(), e.alpha.end());
        frequency.insert(frequency.end(), e.frequency.begin(), e.frequency.end());
        width.insert(width.end(), e.width.begin(), e.width.end());
        siteAlchemical.insert(siteAlchemical.end(), e.siteAlchemical.begin(), e.siteAlchemical.end());
    }
    c->numClusters = (int) clusters.size();
    c->numParticles = particleOffsets.back();
    ixh_pair_desc pd = s->pairTemplate;
    pd.num_particles = c->numParticles;
    pd.alchemical = alchemical.data();
    pd.num_exclusions = (int) exclusions.size()/2;
    pd.exclusions = exclusions.data();
    pd.dispersion = (pd.terms & IXH_TERM_DISPERSION) ? dispersion.data() : nullptr;
    pd.exchange = (pd.terms & IXH_TERM_EXCHANGE) ? exchange.data() : nullptr;
    pd.short_range = (pd.terms & IXH_TERM_SHORT_RANGE) ? shortRange.data() : nullptr;
    pd.default_box = nullptr;
    pd.precision = s->precision;
    ixh_threebody_desc td = s->threebodyTemplate;
    td.num_particles = c->numParticles;
    td.num_sites = siteOffsets.back();
    td.site_offsets = siteParticleOffsets.data();
    td.site_particles = siteParticles.data();
    td.alpha = alpha.data();
    td.frequency = frequency.data();
    td.width = width.data();
    td.molecule = molecule.data();
    td.alchemical = siteAlchemical.data();
    td.precision = s->precision;
    const long long before = deviceBytes.load();
    check(ixh_state_create(s->device, c->numParticles, s->precision, &c->state));
    check(ixh_state_set_stream(c->state, stream));
    check(ixh_state_get_buffers(c->state, &c->buffers));
    if (pd.terms != 0)
        check(ixh_pair_create_segmented(s->device, &pd, c->numClusters, particleOffsets.data(), &c->pair));
    if (s->haveThreebody && td.num_sites > 0)
        check(ixh_threebody_create_segmented(s->device, &td, c->numClusters, siteOffsets.data(), &c->threebody));
    c->energies.reserve(c->numClusters);
    // one evaluation of every cluster, with its particles on a 0.3 nm cubic lattice, loads the kernels and sizes
    // every buffer of the handles, so that no request pays for it
    std::vector<double> lattice(3*(size_t) std::max(c->numParticles, 1), 0.0);
    for (size_t k = 0; k < clusters.size(); k++) {
        const int n = clusters[k]->numParticles;
        int side = 1;
        while (side*side*side < n)
            side++;
        for (int i = 0; i < n; i++) {
            double* x = &lattice[3*(size_t) (particleOffsets[k]+i)];
            x[0] = 0.3*(i%side);
            x[1] = 0.3*((i/side)%side);
            x[2] = 0.3*(i/(side*side));
        }
    }
    check(ixh_state_set_positions(c->state, lattice.data()));
    check(ixh_state_clear(c->state));
    std::vector<int> all(c->numClusters);
    for (int k = 0; k < c->numClusters; k++)
        all[k] = k;
    const int flags = IXH_INCLUDE_FORCES | IXH_INCLUDE_ENERGY;
    hipLaunchKernelGGL(k_zero_doubles, dim3((c->numClusters+255)/256), dim3(256), 0, stream, c->energies.ptr, c->numClusters);
    IXH_HIP_CHECK(hipGetLastError());
    if (c->pair != nullptr)
        check(ixh_pair_execute_subset(c->pair, &c->buffers, s->pairTemplate.terms, flags, c->energies.ptr, all.data(), c->numClusters));
    if (c->threebody != nullptr)
        check(ixh_threebody_execute_subset(c->threebody, &c->buffers, flags, c->energies.ptr, all.data(), c->numClusters));
    IXH_HIP_CHECK(hipStreamSynchronize(stream));
    c->bytes = deviceBytes.load()-before;
    return c;
}

void fail_request(ixh_server* s, const std::shared_ptr<Request>& r, int status, const std::string& message) {
    r->status = status;
    r->error = message;
    r->completion = Clock::now();
    r->done = true;
}

// Launch a batch of cluster requests on a lane: rounds of at most one request per system, one after another on the
// lane's stream, all coordinates uploaded and all results downloaded in one copy each.
void launch_batch(ixh_server* s, Lane& lane, std::vector<std::shared_ptr<Request>>& batch) {
    Catalog& c = *lane.catalog;
    // rounds: the k-th request of a system in the batch goes to round k
    std::map<int, int> seen;
    std::vector<int> round(batch.size());
    int numRounds = 0;
    for (size_t k = 0; k < batch.size(); k++) {
        round[k] = seen[batch[k]->entry]++;
        numRounds = std::max(numRounds, round[k]+1);
    }
    std::vector<std::shared_ptr<Request>> ordered;
    std::vector<int> roundStart(numRounds+1, 0);
    for (int r = 0; r < numRounds; r++) {
        roundStart[r] = (int) ordered.size();
        for (size_t k = 0; k < batch.size(); k++)
            if (round[k] == r)
                ordered.push_back(batch[k]);
    }
    roundStart[numRounds] = (int) ordered.size();
    size_t particles = 0;
    for (auto& r : ordered)
        particles += s->entries[r->entry]->numParticles;
    const int numItems = (int) ordered.size();
    const long long before = deviceBytes.load();
    ensure_capacity(lane, particles, numItems);
    const long long grown = deviceBytes.load()-before;   // staging grown for a cluster registered while running
    int offset = 0;
    for (int k = 0; k < numItems; k++) {
        const Entry& e = *s->entries[ordered[k]->entry];
        const double* p = ordered[k]->positions.data();
        for (int i = 0; i < e.numParticles; i++) {
            float hx, hy, hz, lx, ly, lz;
            split(p[3*i], hx, lx);
            split(p[3*i+1], hy, ly);
            split(p[3*i+2], hz, lz);
            lane.hostPacked[offset+i] = make_float4(hx, hy, hz, 0.0f);
            lane.hostPackedCorrection[offset+i] = make_float4(lx, ly, lz, 0.0f);
        }
        lane.hostItems[k] = make_int3(e.firstParticle, e.numParticles, offset);
        lane.hostSegments[k] = e.segment;
        offset += e.numParticles;
    }
    hipStream_t st = lane.stream;
    const bool mixed = (s->precision == IXH_MIXED);
    const int flags = IXH_INCLUDE_FORCES | IXH_INCLUDE_ENERGY;
    const int energyOffset = 3*(int) particles;
    for (int r = 0; r < numRounds; r++) {
        const int first = roundStart[r], count = roundStart[r+1]-first;
        hipLaunchKernelGGL(k_scatter_positions, dim3(4, count), dim3(256), 0, st, (const float4*) lane.devPacked,
                           (const float4*) lane.devPackedCorrection, (const int3*) lane.devItems+first, count,
                           (float4*) c.buffers.posq, mixed ? (float4*) c.buffers.posq_correction : (float4*) nullptr);
        IXH_HIP_CHECK(hipGetLastError());
        check(ixh_state_clear(c.state));
        hipLaunchKernelGGL(k_zero_doubles, dim3((c.numClusters+255)/256), dim3(256), 0, st, c.energies.ptr, c.numClusters);
        IXH_HIP_CHECK(hipGetLastError());
        if (c.pair != nullptr)
            check(ixh_pair_execute_subset(c.pair, &c.buffers, s->pairTemplate.terms, flags, c.energies.ptr, lane.hostSegments+first,
                                          count));
        if (c.threebody != nullptr)
            check(ixh_threebody_execute_subset(c.threebody, &c.buffers, flags, c.energies.ptr, lane.hostSegments+first, count));
        hipLaunchKernelGGL(k_gather_results, dim3(4, count), dim3(256), 0, st, (const long long*) c.buffers.force,
                           c.buffers.force_stride, (const int3*) lane.devItems+first, (const int*) lane.devSegments+first, count,
                           (const double*) c.energies.ptr, lane.devResults, energyOffset+first);
        IXH_HIP_CHECK(hipGetLastError());
    }
    IXH_HIP_CHECK(hipEventRecord(lane.done, st));
    lane.batch = ordered;
    lane.batchParticles = (int) particles;
    lane.busy = true;
    std::lock_guard<std::mutex> lock(s->mutex);
    s->residentBytes += grown;
    s->batches++;
    s->batchedRequests += numItems;
    s->rounds += numRounds;
}

void finish_batch(ixh_server* s, Lane& lane) {
    const Clock::time_point now = Clock::now();
    int offset = 0;
    const int energyOffset = 3*lane.batchParticles;
    for (size_t k = 0; k < lane.batch.size(); k++) {
        auto& r = lane.batch[k];
        const Entry& e = *s->entries[r->entry];
        r->forces.assign(lane.hostResults+3*offset, lane.hostResults+3*(offset+e.numParticles));
        r->energy = lane.hostResults[energyOffset+k];
        r->completion = now;
        r->done = true;
        s->inFlightBytes -= r->bytes;
        offset += e.numParticles;
    }
    lane.batch.clear();
    lane.busy = false;
}

// Upload a periodic request's coordinates to its slot.
void upload_periodic(ixh_server* s, PeriodicSlot& slot, const Request& r) {
    const Entry& e = *s->entries[slot.entry];
    const int n = e.numParticles;
    const double* p = r.positions.data();
    for (int i = 0; i < n; i++) {
        float hx, hy, hz, lx, ly, lz;
        split(p[3*i], hx, lx);
        split(p[3*i+1], hy, ly);
        split(p[3*i+2], hz, lz);
        slot.hostPosq[i] = make_float4(hx, hy, hz, 0.0f);
        slot.hostCorrection[i] = make_float4(lx, ly, lz, 0.0f);
    }
    hipStream_t st = (hipStream_t) slot.buffers.stream;
    hipLaunchKernelGGL(k_copy_positions, dim3(std::max(1, std::min(64, (n+255)/256))), dim3(256), 0, st, (const float4*) slot.devPosq,
                       (const float4*) slot.devCorrection, n, (float4*) slot.buffers.posq,
                       s->precision == IXH_MIXED ? (float4*) slot.buffers.posq_correction : (float4*) nullptr);
    IXH_HIP_CHECK(hipGetLastError());
}

// Evaluate a periodic request on a slot, asynchronously; the neighbour list is checked on the device.
void launch_periodic(ixh_server* s, PeriodicSlot& slot, const std::shared_ptr<Request>& r, bool deferCheck) {
    const Entry& e = *s->entries[slot.entry];
    const int n = e.numParticles;
    upload_periodic(s, slot, *r);
    hipStream_t st = (hipStream_t) slot.buffers.stream;
    ixh_buffers b = slot.buffers;
    b.lambda = r->lambda;
    const int defer = (deferCheck ? IXH_DEFER_LIST_CHECK : 0);
    if (!r->lambdas.empty()) {
        // the energies of the configuration at every requested lambda: one evaluation plus the alchemical pairs at
        // every value (the fused lambda-window evaluation)
        pair_lambda_launch(slot.pair, &b, r->lambdas.data(), (int) r->lambdas.size(), defer);
    }
    else {
        check(ixh_state_clear(slot.state));
        check(ixh_pair_execute(slot.pair, &b, 7, IXH_INCLUDE_FORCES | IXH_INCLUDE_ENERGY | defer));
        hipLaunchKernelGGL(k_periodic_results, dim3(std::max(1, std::min(64, (n+255)/256))), dim3(256), 0, st,
                           (const long long*) b.force, b.force_stride, n, (const double*) b.energy, slot.devResults);
        IXH_HIP_CHECK(hipGetLastError());
    }
    IXH_HIP_CHECK(hipEventRecord(slot.done, st));
    slot.request = r;
}

void finish_periodic(ixh_server* s, PeriodicSlot& slot) {
    auto r = slot.request;
    const Entry& e = *s->entries[slot.entry];
    int valid = 1;
    check(ixh_pair_list_was_valid(slot.pair, &valid));
    if (!valid) {
        // the coordinates moved beyond the list's skin: evaluate again, rebuilding the list
        launch_periodic(s, slot, r, false);
        IXH_HIP_CHECK(hipEventSynchronize(slot.done));
    }
    const int n = e.numParticles;
    if (!r->lambdas.empty()) {
        r->energies.resize(r->lambdas.size());
        pair_lambda_collect(slot.pair, (int) r->lambdas.size(), r->energies.data());
        for (size_t k = 0; k < r->lambdas.size(); k++) {
            double lrc = 0;
            check(ixh_pair_long_range_energy(slot.pair, e.box, r->lambdas[k], &lrc));
            r->energies[k] += lrc;
        }
        r->energy = r->energies[0];
    }
    else {
        double lrc = 0;
        check(ixh_pair_long_range_energy(slot.pair, e.box, r->lambda, &lrc));
        r->forces.assign(slot.hostResults, slot.hostResults+3*(size_t) n);
        r->energy = slot.hostResults[3*(size_t) n]+lrc;
    }
    r->completion = Clock::now();
    r->done = true;
    s->inFlightBytes -= r->bytes;
    slot.request.reset();
}

// Whether a request fits the memory budget beside those in flight (always, when none is).
bool fits(const ixh_server* s, long long bytes) {
    return s->inFlightBytes == 0 || s->inFlightBytes+bytes <= s->memoryBudget;
}

void admitted(ixh_server* s, long long bytes) {
    s->inFlightBytes += bytes;
    s->peakInFlightBytes = std::max(s->peakInFlightBytes, s->inFlightBytes);
}

void serve(ixh_server* s) {
    IXH_HIP_CHECK(hipSetDevice(s->device));
    std::deque<std::shared_ptr<Request>> pending;
    std::vector<std::unique_ptr<Catalog>> retired;
    while (true) {
        int slots = 0;
        {
            std::unique_lock<std::mutex> lock(s->mutex);
            bool anyBusy = s->lanes[0].busy || s->lanes[1].busy;
            slots = s->numSlots.load();
            for (int k = 0; k < slots; k++)
                anyBusy = anyBusy || s->periodicSlots[k]->request != nullptr;
            if (!s->stopping && s->incoming.empty() && pending.empty() && !anyBusy)
                s->submitted.wait(lock, [&] { return s->stopping || !s->incoming.empty(); });
            if (s->stopping && s->incoming.empty() && pending.empty() && !anyBusy)
                return;
            while (!s->incoming.empty()) {
                pending.push_back(s->incoming.front());
                s->incoming.pop_front();
            }
            slots = s->numSlots.load();
            // an idle lane takes up a newer catalog (the old one is released below, outside the lock)
            for (Lane& lane : s->lanes)
                if (!lane.busy && lane.pending != nullptr) {
                    if (lane.catalog != nullptr) {
                        s->residentBytes -= lane.catalog->bytes;
                        retired.push_back(std::move(lane.catalog));
                    }
                    lane.catalog = std::move(lane.pending);
                }
        }
        retired.clear();
        bool progress = false;
        // completions
        for (Lane& lane : s->lanes)
            if (lane.busy && hipEventQuery(lane.done) == hipSuccess) {
                std::lock_guard<std::mutex> lock(s->mutex);
                finish_batch(s, lane);
                progress = true;
            }
        for (int k = 0; k < slots; k++) {
            PeriodicSlot& slot = *s->periodicSlots[k];
            if (slot.request != nullptr && hipEventQuery(slot.done) == hipSuccess) {
                std::lock_guard<std::mutex> lock(s->mutex);
                try {
                    finish_periodic(s, slot);
                }
                catch (const std::exception& ex) {
                    s->inFlightBytes -= slot.request->bytes;
                    fail_request(s, slot.request, IXH_ERROR_HIP, ex.what());
                    slot.request.reset();
                }
                progress = true;
            }
        }
        if (progress)
            s->completed.notify_all();
        // admissions: every free lane takes the pending cluster requests of its catalog that fit, in arrival order
        for (Lane& lane : s->lanes) {
            if (lane.busy || lane.catalog == nullptr)
                continue;
            std::vector<std::shared_ptr<Request>> batch;
            long long bytes = 0;
            for (auto it = pending.begin(); it != pending.end() && (int) batch.size() < s->maxBatch;) {
                const Entry& e = *s->entries[(*it)->entry];
                if (!e.periodic && e.segment < lane.catalog->numClusters &&
                    ((s->inFlightBytes == 0 && batch.empty()) || s->inFlightBytes+bytes+(*it)->bytes <= s->memoryBudget)) {
                    bytes += (*it)->bytes;
                    batch.push_back(*it);
                    it = pending.erase(it);
                }
                else
                    ++it;
            }
            if (!batch.empty()) {
                admitted(s, bytes);
                try {
                    launch_batch(s, lane, batch);
                }
                catch (const std::exception& ex) {
                    std::lock_guard<std::mutex> lock(s->mutex);
                    s->inFlightBytes -= bytes;
                    for (auto& r : batch)
                        fail_request(s, r, IXH_ERROR_HIP, ex.what());
                    s->completed.notify_all();
                }
                progress = true;
            }
        }
        // periodic requests: each takes a free slot of its system, in arrival order, within the budget
        for (auto it = pending.begin(); it != pending.end();) {
            const Entry& e = *s->entries[(*it)->entry];
            PeriodicSlot* free = nullptr;
            if (e.periodic && fits(s, (*it)->bytes))
                for (int k : e.slots)
                    if (k < slots && s->periodicSlots[k]->request == nullptr) {
                        free = s->periodicSlots[k].get();
                        break;
                    }
            if (free == nullptr) {
                ++it;
                continue;
            }
            auto r = *it;
            it = pending.erase(it);
            admitted(s, r->bytes);
            try {
                launch_periodic(s, *free, r, true);
            }
            catch (const std::exception& ex) {
                std::lock_guard<std::mutex> lock(s->mutex);
                free->request.reset();
                s->inFlightBytes -= r->bytes;
                fail_request(s, r, IXH_ERROR_HIP, ex.what());
                s->completed.notify_all();
            }
            progress = true;
        }
        if (!progress)
            std::this_thread::sleep_for(std::chrono::microseconds(5));
    }
}

void copy_pair_entry(Entry& e, const ixh_pair_desc* d) {
    const int n = d->num_particles;
    e.numParticles = n;
    if (d->terms & IXH_TERM_DISPERSION)
        e.dispersion.assign(d->dispersion, d->dispersion+5*(size_t) n);
    if (d->terms & IXH_TERM_EXCHANGE)
        e.exchange.assign(d->exchange, d->exchange+2*(size_t) n);
    if (d->terms & IXH_TERM_SHORT_RANGE)
        e.shortRange.assign(d->short_range, d->short_range+3*(size_t) n);
    e.alchemical.assign(n, 0);
    if (d->alchemical != nullptr)
        e.alchemical.assign(d->alchemical, d->alchemical+n);
    e.exclusions.assign(d->exclusions, d->exclusions+2*(size_t) d->num_exclusions);
}

bool same_pair_settings(const ixh_pair_desc& a, const ixh_pair_desc& b) {
    return a.terms == b.terms && a.method == b.method && a.softcore_radius == b.softcore_radius && a.damping_scale == b.damping_scale &&
           a.exchange_overlap_method == b.exchange_overlap_method && a.short_range_overlap_method == b.short_range_overlap_method;
}

bool same_threebody_settings(const ixh_threebody_desc& a, const ixh_threebody_desc& b) {
    return a.cutoff == b.cutoff && a.scale_factor == b.scale_factor && a.damping_method == b.damping_method &&
           a.damping_scale == b.damping_scale && a.centroid_damping_radius == b.centroid_damping_radius && a.periodic == b.periodic;
}

// The execution slots of a periodic system, their neighbour lists built (synchronous).
void create_slots(ixh_server* s, Entry& e, int index, const ixh_pair_desc* pair, const double* box, const double* positions) {
    ixh_pair_desc d = *pair;
    d.precision = s->precision;
    d.default_box = box;
    const int n = e.numParticles;
    for (int k = 0; k < s->slotsPerPeriodic; k++) {
        const long long before = deviceBytes.load();
        std::unique_ptr<PeriodicSlot> slot(new PeriodicSlot());
        slot->entry = index;
        check(ixh_pair_create(s->device, &d, &slot->pair));
        check(ixh_state_create(s->device, n, s->precision, &slot->state));
        check(ixh_state_set_stream(slot->state, s->periodicStream));
        check(ixh_state_set_box(slot->state, box));
        check(ixh_state_get_buffers(slot->state, &slot->buffers));
        mapped_alloc(slot->hostPosq, slot->devPosq, (size_t) n);
        mapped_alloc(slot->hostCorrection, slot->devCorrection, (size_t) n);
        mapped_alloc(slot->hostResults, slot->devResults, 3*(size_t) n+1);
        IXH_HIP_CHECK(hipEventCreateWithFlags(&slot->done, hipEventDisableTiming));
        if (positions != nullptr) {
            // the first evaluation builds the neighbour list and sizes the handle's buffers, including those of a
            // window request
            check(ixh_state_set_positions(slot->state, positions));
            ixh_buffers b = slot->buffers;
            check(ixh_pair_execute(slot->pair, &b, 7, IXH_INCLUDE_FORCES | IXH_INCLUDE_ENERGY));
            const double one = 1.0;
            pair_lambda_launch(slot->pair, &b, &one, 1, 0);
        }
        IXH_HIP_CHECK(hipStreamSynchronize(s->periodicStream));
        e.slotBytes = std::max(e.slotBytes, deviceBytes.load()-before);
        e.slots.push_back((int) s->periodicSlots.size());
        s->periodicSlots.push_back(std::move(slot));
    }
}

// Publish a registered entry (and the slots created for it).
int publish(ixh_server* s, std::unique_ptr<Entry> e) {
    std::lock_guard<std::mutex> lock(s->mutex);
    const int index = (int) s->entries.size();
    s->entries.push_back(std::move(e));
    s->numEntries.store(index+1);
    s->numSlots.store((int) s->periodicSlots.size());
    return index;
}

}  // namespace

extern "C" {

int ixh_server_create(int device, int precision, int max_batch, long long memory_budget, ixh_server** out) {
    return guarded([&] {
        if (out == nullptr || max_batch <= 0 || memory_budget <= 0)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_create: invalid argument");
        if (precision != IXH_SINGLE && precision != IXH_MIXED)
            throw Error(IXH_ERROR_UNSUPPORTED, "ixh_server_create: the server runs in single or mixed precision");
        DeviceGuard guard(device);
        std::unique_ptr<ixh_server> s(new ixh_server());
        s->device = device;
        s->precision = precision;
        s->maxBatch = max_batch;
        s->memoryBudget = memory_budget;
        s->entries.reserve(MAX_SYSTEMS);
        s->periodicSlots.reserve(MAX_SLOTS);
        for (Lane& lane : s->lanes) {
            IXH_HIP_CHECK(hipStreamCreateWithFlags(&lane.stream, hipStreamNonBlocking));
            IXH_HIP_CHECK(hipEventCreateWithFlags(&lane.done, hipEventDisableTiming));
        }
        IXH_HIP_CHECK(hipStreamCreateWithFlags(&s->periodicStream, hipStreamNonBlocking));
        *out = s.release();
    });
}

int ixh_server_add_cluster(ixh_server* s, const ixh_pair_desc* pair, const ixh_threebody_desc* threebody, int* entry) {
    return guarded([&] {
        if (s == nullptr || pair == nullptr || entry == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_add_cluster: null argument");
        if (pair->method != IXH_NO_CUTOFF)
            throw Error(IXH_ERROR_UNSUPPORTED, "ixh_server_add_cluster: cluster systems use the NoCutoff method");
        std::lock_guard<std::mutex> add(s->addMutex);
        if ((int) s->entries.size() >= MAX_SYSTEMS)
            throw Error(IXH_ERROR_CAPACITY, "ixh_server_add_cluster: too many systems");
        if (s->haveClusterTemplate && (!same_pair_settings(s->pairTemplate, *pair) || s->haveThreebody != (threebody != nullptr) ||
                 (threebody != nullptr && !same_threebody_settings(s->threebodyTemplate, *threebody))))
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_add_cluster: every cluster must have the same force settings");
        if (threebody != nullptr && (threebody->periodic || threebody->num_particles != pair->num_particles))
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_add_cluster: the three-body force must be nonperiodic over the same particles");
        std::unique_ptr<Entry> e(new Entry());
        copy_pair_entry(*e, pair);
        if (threebody != nullptr) {
            const int ns = threebody->num_sites;
            e->numSites = ns;
            e->siteOffsets.assign(threebody->site_offsets, threebody->site_offsets+ns+1);
            e->siteParticles.assign(threebody->site_particles, threebody->site_particles+e->siteOffsets[ns]);
            e->alpha.assign(threebody->alpha, threebody->alpha+ns);
            e->frequency.assign(threebody->frequency, threebody->frequency+ns);
            e->width.assign(threebody->width, threebody->width+ns);
            e->molecule.assign(threebody->molecule, threebody->molecule+ns);
            e->siteAlchemical.assign(ns, 0);
            if (threebody->alchemical != nullptr)
                e->siteAlchemical.assign(threebody->alchemical, threebody->alchemical+ns);
        }
        if (!s->haveClusterTemplate) {
            s->pairTemplate = *pair;
            s->haveThreebody = (threebody != nullptr);
            if (threebody != nullptr)
                s->threebodyTemplate = *threebody;
            s->haveClusterTemplate = true;
        }
        // clusters are appended to the catalog: the earlier ones keep their segments and particles
        e->segment = s->numClusters;
        e->firstParticle = s->catalogParticles;
        const int n = e->numParticles;
        if (!s->started) {
            *entry = publish(s, std::move(e));
            s->numClusters++;
            s->catalogParticles += n;
            return;
        }
        // running: a new catalog for both lanes, built here while the lanes serve the old one; the entry is published
        // with it
        DeviceGuard guard(s->device);
        const int count = (int) s->entries.size();
        s->entries.push_back(std::move(e));   // visible to this thread only until published (numEntries)
        std::unique_ptr<Catalog> next[2];
        try {
            for (int k = 0; k < 2; k++)
                next[k] = build_catalog(s, count+1, s->lanes[k].stream);
        }
        catch (...) {
            s->entries.pop_back();
            throw;
        }
        {
            std::lock_guard<std::mutex> lock(s->mutex);
            for (int k = 0; k < 2; k++) {
                s->residentBytes += next[k]->bytes-(s->lanes[k].pending != nullptr ? s->lanes[k].pending->bytes : 0);
                s->lanes[k].pending = std::move(next[k]);
            }
            s->numEntries.store(count+1);
        }
        s->numClusters++;
        s->catalogParticles += n;
        *entry = count;
        s->submitted.notify_all();
    });
}

int ixh_server_add_periodic(ixh_server* s, const ixh_pair_desc* pair, const double* box, const double* positions, int* entry) {
    return guarded([&] {
        if (s == nullptr || pair == nullptr || box == nullptr || entry == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_add_periodic: null argument");
        std::lock_guard<std::mutex> add(s->addMutex);
        if ((int) s->entries.size() >= MAX_SYSTEMS || (int) s->periodicSlots.size()+s->slotsPerPeriodic > MAX_SLOTS)
            throw Error(IXH_ERROR_CAPACITY, "ixh_server_add_periodic: too many systems");
        DeviceGuard guard(s->device);
        std::unique_ptr<Entry> e(new Entry());
        e->periodic = true;
        e->numParticles = pair->num_particles;
        std::copy(box, box+9, e->box);
        const int index = (int) s->entries.size();
        const size_t firstSlot = s->periodicSlots.size();
        try {
            create_slots(s, *e, index, pair, box, positions);
        }
        catch (...) {
            std::lock_guard<std::mutex> lock(s->mutex);
            while (s->periodicSlots.size() > firstSlot)
                s->periodicSlots.pop_back();
            throw;
        }
        {
            std::lock_guard<std::mutex> lock(s->mutex);
            s->residentBytes += e->slotBytes*s->slotsPerPeriodic;
        }
        *entry = publish(s, std::move(e));
        s->submitted.notify_all();
    });
}

int ixh_server_start(ixh_server* s) {
    return guarded([&] {
        if (s == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_start: null argument");
        std::lock_guard<std::mutex> add(s->addMutex);
        if (s->started)
            return;
        DeviceGuard guard(s->device);
        if (s->numClusters > 0) {
            int largest = 0;
            for (int k = 0; k < s->numEntries.load(); k++)
                if (!s->entries[k]->periodic)
                    largest = std::max(largest, s->entries[k]->numParticles);
            for (Lane& lane : s->lanes) {
                lane.catalog = build_catalog(s, s->numEntries.load(), lane.stream);
                s->residentBytes += lane.catalog->bytes;
                // staging for the largest batch (max_batch requests of the largest cluster), allocated now
                const long long before = deviceBytes.load();
                ensure_capacity(lane, (size_t) s->maxBatch*largest, (size_t) s->maxBatch);
                s->residentBytes += deviceBytes.load()-before;
            }
        }
        s->started = true;
        s->worker = std::thread([s] {
            try {
                serve(s);
            }
            catch (const std::exception& ex) {
                std::lock_guard<std::mutex> lock(s->mutex);
                s->fatal = ex.what();
                for (auto& t : s->tickets)
                    if (!t.second->done)
                        fail_request(s, t.second, IXH_ERROR_HIP, ex.what());
                s->completed.notify_all();
            }
        });
    });
}

}  // extern "C"

namespace {

long long submit(ixh_server* s, int entry, const double* positions, double lambda, const double* lambdas, int count) {
    if (s == nullptr || positions == nullptr)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_submit: null argument");
    if (!s->started)
        throw Error(IXH_ERROR_STATE, "ixh_server_submit: the server has not started");
    if (entry < 0 || entry >= s->numEntries.load())
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_submit: unknown system");
    const Entry& e = *s->entries[entry];
    auto r = std::make_shared<Request>();
    r->entry = entry;
    r->lambda = lambda;
    if (count > 0) {
        if (!e.periodic)
            throw Error(IXH_ERROR_UNSUPPORTED, "ixh_server_submit_windows: window requests are served for periodic systems");
        if (lambdas == nullptr || count > 16)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_submit_windows: between 1 and 16 lambda values");
        r->lambdas.assign(lambdas, lambdas+count);
    }
    r->positions.assign(positions, positions+3*(size_t) e.numParticles);
    r->bytes = (e.periodic ? e.slotBytes : cluster_request_bytes(e.numParticles));
    r->arrival = Clock::now();
    {
        std::lock_guard<std::mutex> lock(s->mutex);
        if (!s->fatal.empty())
            throw Error(IXH_ERROR_HIP, "ixh_server_submit: the server has stopped: "+s->fatal);
        r->ticket = s->nextTicket++;
        s->tickets[r->ticket] = r;
        s->incoming.push_back(r);
    }
    s->submitted.notify_one();
    return r->ticket;
}

std::shared_ptr<Request> wait(ixh_server* s, long long ticket) {
    if (s == nullptr)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_wait: null argument");
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lock(s->mutex);
        auto it = s->tickets.find(ticket);
        if (it == s->tickets.end())
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_wait: unknown ticket");
        r = it->second;
        s->completed.wait(lock, [&] { return r->done; });
        s->tickets.erase(it);
    }
    if (r->status != IXH_OK)
        throw Error(r->status, r->error);
    return r;
}

double latency_ms(const Request& r) {
    return std::chrono::duration<double, std::milli>(r.completion-r.arrival).count();
}

}  // namespace

extern "C" {

int ixh_server_submit(ixh_server* s, int entry, const double* positions, double lambda, long long* ticket) {
    return guarded([&] {
        if (ticket == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_submit: null argument");
        *ticket = submit(s, entry, positions, lambda, nullptr, 0);
    });
}

int ixh_server_submit_windows(ixh_server* s, int entry, const double* positions, const double* lambdas, int count, long long* ticket) {
    return guarded([&] {
        if (ticket == nullptr || count < 1)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_submit_windows: invalid argument");
        *ticket = submit(s, entry, positions, 1.0, lambdas, count);
    });
}

int ixh_server_wait(ixh_server* s, long long ticket, double* energy, double* forces, double* latency) {
    return guarded([&] {
        auto r = wait(s, ticket);
        if (!r->lambdas.empty())
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_wait: a window request (ixh_server_wait_windows)");
        if (energy != nullptr)
            *energy = r->energy;
        if (forces != nullptr)
            std::copy(r->forces.begin(), r->forces.end(), forces);
        if (latency != nullptr)
            *latency = latency_ms(*r);
    });
}

int ixh_server_wait_windows(ixh_server* s, long long ticket, double* energies, double* latency) {
    return guarded([&] {
        auto r = wait(s, ticket);
        if (r->lambdas.empty())
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_wait_windows: not a window request");
        if (energies != nullptr)
            std::copy(r->energies.begin(), r->energies.end(), energies);
        if (latency != nullptr)
            *latency = latency_ms(*r);
    });
}

int ixh_server_statistics(ixh_server* s, long long* batches, long long* requests) {
    return guarded([&] {
        if (s == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_statistics: null argument");
        std::lock_guard<std::mutex> lock(s->mutex);
        if (batches != nullptr)
            *batches = s->batches;
        if (requests != nullptr)
            *requests = s->batchedRequests;
    });
}

int ixh_server_memory(ixh_server* s, long long* resident, long long* in_flight, long long* peak_in_flight) {
    return guarded([&] {
        if (s == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_memory: null argument");
        std::lock_guard<std::mutex> lock(s->mutex);
        if (resident != nullptr)
            *resident = s->residentBytes;
        if (in_flight != nullptr)
            *in_flight = s->inFlightBytes;
        if (peak_in_flight != nullptr)
            *peak_in_flight = s->peakInFlightBytes;
    });
}

int ixh_server_replay(ixh_server* s, int n, const double* arrival, const int* entries, const double* lambdas, const long long* offsets,
                      const double* positions, const int* window_counts, const double* window_lambdas, double* latency,
                      double* energies, double* makespan_s) {
    return guarded([&] {
        if (s == nullptr || n < 0 || (n > 0 && (arrival == nullptr || entries == nullptr || offsets == nullptr || positions == nullptr)))
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_replay: null argument");
        if (window_counts != nullptr && window_lambdas == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_replay: window counts without window values");
        std::vector<long long> tickets(n);
        const Clock::time_point start = Clock::now();
        // submit at the scheduled times (the latency clock of a request starts at its scheduled arrival)
        long long windowOffset = 0;
        for (int k = 0; k < n; k++) {
            const Clock::time_point due = start+std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival[k]));
            while (Clock::now() < due) {
            }
            const int count = (window_counts != nullptr ? window_counts[k] : 0);
            tickets[k] = submit(s, entries[k], positions+offsets[k], lambdas != nullptr ? lambdas[k] : 1.0,
                                count > 0 ? window_lambdas+windowOffset : nullptr, count);
            windowOffset += count;
            std::lock_guard<std::mutex> lock(s->mutex);
            s->tickets[tickets[k]]->arrival = due;
        }
        Clock::time_point last = start;
        for (int k = 0; k < n; k++) {
            auto r = wait(s, tickets[k]);
            if (latency != nullptr)
                latency[k] = latency_ms(*r);
            if (energies != nullptr)
                energies[k] = r->energy;
            last = std::max(last, r->completion);
        }
        if (makespan_s != nullptr)
            *makespan_s = std::chrono::duration<double>(last-start).count();
    });
}

int ixh_server_destroy(ixh_server* s) {
    return guarded([&] {
        if (s == nullptr)
            return;
        {
            std::lock_guard<std::mutex> lock(s->mutex);
            s->stopping = true;
        }
        s->submitted.notify_all();
        if (s->worker.joinable())
            s->worker.join();
        DeviceGuard guard(s->device);
        (void) hipDeviceSynchronize();
        for (Lane& lane : s->lanes) {
            lane.catalog.reset();
            lane.pending.reset();
            release_lane_staging(lane);
            if (lane.done != nullptr)
                (void) hipEventDestroy(lane.done);
        }
        s->periodicSlots.clear();
        for (Lane& lane : s->lanes)
            if (lane.stream != nullptr)
                (void) hipStreamDestroy(lane.stream);
        if (s->periodicStream != nullptr)
            (void) hipStreamDestroy(s->periodicStream);
        delete s;
    });
}

}  // extern "C"
