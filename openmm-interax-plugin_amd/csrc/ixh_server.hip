// A serving loop over device-resident systems (the served configuration: a Poisson stream of cluster evaluations
// and alchemical λ-window evaluations).
//
// No counterpart in the reference.  Systems are registered once; their parameters stay on the device.  Cluster
// systems form one catalog, concatenated into segmented pair and three-body handles, instantiated twice ("lanes",
// each with its own stream and state) so that two batches are on the GPU at once.  The server thread keeps a FIFO of
// pending requests; whenever a lane is free it admits, in arrival order, the pending cluster requests that fit the
// batch size and the memory budget (bytes of work in flight), at most one per system, uploads their coordinates in
// one copy, evaluates exactly their segments, gathers their forces and energies into one download, and signals the
// requests when the lane's completion event has passed.  Periodic systems (alchemical boxes with a cutoff) have their
// own handle and are evaluated one request at a time on a third stream.  A request's result does not depend on the
// other requests of its batch: segments are evaluated independently and their sums have a fixed order.
#include "ixh_common.h"

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

void check(int status) {
    if (status != IXH_OK)
        throw Error(status, ixh_last_error());
}

// Scatter the coordinates of a batch, packed request after request, into the catalog's position buffers.
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

// Gather the forces (fixed point to kJ/mol/nm) of a batch's requests and their segment energies into one array:
// per request 3 x count forces at 3 x offset, then the energies after all forces.
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

__global__ void k_zero_doubles(double* p, int n) {
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n)
        p[i] = 0;
}

struct Request {
    long long ticket = 0;
    int entry = 0;
    double lambda = 1;
    std::vector<double> positions;
    Clock::time_point arrival;
    Clock::time_point completion;
    bool done = false;
    int status = IXH_OK;
    std::string error;
    double energy = 0;
    std::vector<double> forces;
};

// A registered system: a cluster of the catalog, or a periodic system with its own handle.
struct Entry {
    bool periodic = false;
    int numParticles = 0;
    long long footprint = 0;        // bytes of device work in flight while one of its requests is evaluated
    // cluster: its segment and particle range in the catalog
    int segment = -1, firstParticle = 0;
    int numSites = 0;
    std::vector<double> dispersion, exchange, shortRange, alpha, frequency, width;
    std::vector<unsigned char> alchemical, siteAlchemical;
    std::vector<int> exclusions, siteOffsets, siteParticles, molecule;
    // periodic: the box and the execution slots of the system
    double box[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<int> slots;
    bool inFlight = false;
};

// One execution slot of a periodic system: a handle, a state, staging and results; one request at a time.
struct PeriodicSlot {
    int entry = 0;
    ixh_pair* pair = nullptr;
    ixh_state* state = nullptr;
    ixh_buffers buffers;
    float4* hostPosq = nullptr;          // pinned staging
    float4* hostCorrection = nullptr;
    double* hostResults = nullptr;       // forces (3n) and the energy
    DeviceArray<double> results;
    hipEvent_t done = nullptr;
    std::shared_ptr<Request> request;
};

struct Lane {
    ixh_pair* pair = nullptr;
    ixh_threebody* threebody = nullptr;
    ixh_state* state = nullptr;
    ixh_buffers buffers;
    hipStream_t stream = nullptr;
    DeviceArray<double> energies;       // one per catalog segment
    DeviceArray<float4> packed, packedCorrection;
    DeviceArray<int3> items;
    DeviceArray<int> segments;
    DeviceArray<double> results;
    float4* hostPacked = nullptr;       // pinned staging: positions, corrections, items, segments
    float4* hostPackedCorrection = nullptr;
    int3* hostItems = nullptr;
    int* hostSegments = nullptr;
    double* hostResults = nullptr;
    size_t capacityParticles = 0, capacityItems = 0;
    hipEvent_t done = nullptr;
    std::vector<std::shared_ptr<Request>> batch;
    bool busy = false;
};

}  // namespace

struct ixh_server {
    int device = 0;
    int precision = IXH_MIXED;
    int maxBatch = 64;
    long long memoryBudget = 0;
    std::vector<std::unique_ptr<Entry>> entries;
    int numClusters = 0, catalogParticles = 0;
    // settings shared by the catalog's clusters (from the first registered cluster)
    ixh_pair_desc pairTemplate;
    ixh_threebody_desc threebodyTemplate;
    bool haveThreebody = false;
    Lane lanes[2];
    std::vector<std::unique_ptr<PeriodicSlot>> periodicSlots;
    int slotsPerPeriodic = 4;
    std::thread worker;
    std::mutex mutex;
    std::condition_variable submitted, completed;
    std::deque<std::shared_ptr<Request>> incoming;
    std::map<long long, std::shared_ptr<Request>> tickets;
    long long nextTicket = 1;
    bool started = false, stopping = false;
    long long batches = 0, batchedRequests = 0;
    std::string fatal;
};

namespace {

long long cluster_footprint(const Entry& e) {
    const long long n = e.numParticles, s = e.numSites;
    // positions and corrections (32 B), fixed-point forces (24 B), gathered results (24 B) per particle; the pair
    // tiles of the cluster (64 x 64 masks and energies); the three-body neighbour lists (36 B per entry)
    const long long blocks = (n+63)/64;
    return 80*n+blocks*(blocks+1)/2*(64*8+8)+s*(s-1)/2*36;
}

void ensure_capacity(Lane& lane, size_t particles, size_t items, int numResults) {
    if (particles > lane.capacityParticles) {
        if (lane.hostPacked != nullptr)
            IXH_HIP_CHECK(hipHostFree(lane.hostPacked));
        if (lane.hostPackedCorrection != nullptr)
            IXH_HIP_CHECK(hipHostFree(lane.hostPackedCorrection));
        lane.capacityParticles = std::max<size_t>(particles, 4096);
        IXH_HIP_CHECK(hipHostMalloc((void**) &lane.hostPacked, sizeof(float4)*lane.capacityParticles, hipHostMallocDefault));
        IXH_HIP_CHECK(hipHostMalloc((void**) &lane.hostPackedCorrection, sizeof(float4)*lane.capacityParticles, hipHostMallocDefault));
        lane.packed.reserve(lane.capacityParticles);
        lane.packedCorrection.reserve(lane.capacityParticles);
        if (lane.hostResults != nullptr)
            IXH_HIP_CHECK(hipHostFree(lane.hostResults));
        lane.hostResults = nullptr;
    }
    if (items > lane.capacityItems) {
        if (lane.hostItems != nullptr)
            IXH_HIP_CHECK(hipHostFree(lane.hostItems));
        if (lane.hostSegments != nullptr)
            IXH_HIP_CHECK(hipHostFree(lane.hostSegments));
        lane.capacityItems = std::max<size_t>(items, 256);
        IXH_HIP_CHECK(hipHostMalloc((void**) &lane.hostItems, sizeof(int3)*lane.capacityItems, hipHostMallocDefault));
        IXH_HIP_CHECK(hipHostMalloc((void**) &lane.hostSegments, sizeof(int)*lane.capacityItems, hipHostMallocDefault));
        lane.items.reserve(lane.capacityItems);
        lane.segments.reserve(lane.capacityItems);
        if (lane.hostResults != nullptr)
            IXH_HIP_CHECK(hipHostFree(lane.hostResults));
        lane.hostResults = nullptr;
    }
    if (lane.hostResults == nullptr) {
        size_t n = 3*lane.capacityParticles+lane.capacityItems;
        IXH_HIP_CHECK(hipHostMalloc((void**) &lane.hostResults, sizeof(double)*n, hipHostMallocDefault));
        lane.results.reserve(n);
    }
    (void) numResults;
}

void split(double x, float& hi, float& lo) {
    hi = (float) x;
    lo = (float) (x-(double) hi);
}

// Build the catalog: concatenate the registered clusters and create the segmented handles of both lanes.
void build_catalog(ixh_server* s) {
    std::vector<Entry*> clusters;
    for (auto& e : s->entries)
        if (!e->periodic)
            clusters.push_back(e.get());
    s->numClusters = (int) clusters.size();
    if (clusters.empty())
        return;
    std::vector<int> particleOffsets(clusters.size()+1, 0), siteOffsets(clusters.size()+1, 0);
    std::vector<double> dispersion, exchange, shortRange, alpha, frequency, width;
    std::vector<unsigned char> alchemical, siteAlchemical;
    std::vector<int> exclusions, siteParticleOffsets(1, 0), siteParticles, molecule;
    for (size_t k = 0; k < clusters.size(); k++) {
        Entry& e = *clusters[k];
        e.segment = (int) k;
        e.firstParticle = particleOffsets[k];
        particleOffsets[k+1] = particleOffsets[k]+e.numParticles;
        dispersion.insert(dispersion.end(), e.dispersion.begin(), e.dispersion.end());
        exchange.insert(exchange.end(), e.exchange.begin(), e.exchange.end());
        shortRange.insert(shortRange.end(), e.shortRange.begin(), e.shortRange.end());
        alchemical.insert(alchemical.end(), e.alchemical.begin(), e.alchemical.end());
        for (int x : e.exclusions)
            exclusions.push_back(x+e.firstParticle);
        siteOffsets[k+1] = siteOffsets[k]+e.numSites;
        for (int i = 0; i < e.numSites; i++) {
            for (int q = e.siteOffsets[i]; q < e.siteOffsets[i+1]; q++)
                siteParticles.push_back(e.siteParticles[q]+e.firstParticle);
            siteParticleOffsets.push_back((int) siteParticles.size());
            // negative molecule indices become -1 - (catalog site index)
            molecule.push_back(e.molecule[i] < 0 ? -1-(siteOffsets[k]+i) : e.molecule[i]);
        }
        alpha.insert(alpha.end(), e.alpha.begin(), e.alpha.end());
        frequency.insert(frequency.end(), e.frequency.begin(), e.frequency.end());
        width.insert(width.end(), e.width.begin(), e.width.end());
        siteAlchemical.insert(siteAlchemical.end(), e.siteAlchemical.begin(), e.siteAlchemical.end());
    }
    s->catalogParticles = particleOffsets.back();
    ixh_pair_desc pd = s->pairTemplate;
    pd.num_particles = s->catalogParticles;
    pd.alchemical = alchemical.data();
    pd.num_exclusions = (int) exclusions.size()/2;
    pd.exclusions = exclusions.data();
    pd.dispersion = (pd.terms & IXH_TERM_DISPERSION) ? dispersion.data() : nullptr;
    pd.exchange = (pd.terms & IXH_TERM_EXCHANGE) ? exchange.data() : nullptr;
    pd.short_range = (pd.terms & IXH_TERM_SHORT_RANGE) ? shortRange.data() : nullptr;
    pd.default_box = nullptr;
    pd.precision = s->precision;
    ixh_threebody_desc td = s->threebodyTemplate;
    td.num_particles = s->catalogParticles;
    td.num_sites = siteOffsets.back();
    td.site_offsets = siteParticleOffsets.data();
    td.site_particles = siteParticles.data();
    td.alpha = alpha.data();
    td.frequency = frequency.data();
    td.width = width.data();
    td.molecule = molecule.data();
    td.alchemical = siteAlchemical.data();
    td.precision = s->precision;
    for (Lane& lane : s->lanes) {
        check(ixh_state_create(s->device, s->catalogParticles, s->precision, &lane.state));
        check(ixh_state_get_buffers(lane.state, &lane.buffers));
        lane.stream = (hipStream_t) lane.buffers.stream;
        if (pd.terms != 0)
            check(ixh_pair_create_segmented(s->device, &pd, s->numClusters, particleOffsets.data(), &lane.pair));
        if (s->haveThreebody && td.num_sites > 0)
            check(ixh_threebody_create_segmented(s->device, &td, s->numClusters, siteOffsets.data(), &lane.threebody));
        lane.energies.reserve(s->numClusters);
        IXH_HIP_CHECK(hipEventCreateWithFlags(&lane.done, hipEventDisableTiming));
    }
}

void fail_request(ixh_server* s, const std::shared_ptr<Request>& r, const std::string& message) {
    r->status = IXH_ERROR_HIP;
    r->error = message;
    r->completion = Clock::now();
    r->done = true;
}

// Launch a batch of cluster requests on a lane.
void launch_batch(ixh_server* s, Lane& lane, std::vector<std::shared_ptr<Request>>& batch) {
    size_t particles = 0;
    for (auto& r : batch)
        particles += s->entries[r->entry]->numParticles;
    const int numItems = (int) batch.size();
    ensure_capacity(lane, particles, numItems, numItems);
    int offset = 0;
    for (int k = 0; k < numItems; k++) {
        const Entry& e = *s->entries[batch[k]->entry];
        const double* p = batch[k]->positions.data();
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
    IXH_HIP_CHECK(hipMemcpyAsync(lane.packed.ptr, lane.hostPacked, sizeof(float4)*particles, hipMemcpyHostToDevice, st));
    const bool mixed = (s->precision == IXH_MIXED);
    if (mixed)
        IXH_HIP_CHECK(hipMemcpyAsync(lane.packedCorrection.ptr, lane.hostPackedCorrection, sizeof(float4)*particles,
                                     hipMemcpyHostToDevice, st));
    IXH_HIP_CHECK(hipMemcpyAsync(lane.items.ptr, lane.hostItems, sizeof(int3)*numItems, hipMemcpyHostToDevice, st));
    IXH_HIP_CHECK(hipMemcpyAsync(lane.segments.ptr, lane.hostSegments, sizeof(int)*numItems, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_scatter_positions, dim3(4, numItems), dim3(256), 0, st, (const float4*) lane.packed.ptr,
                       (const float4*) lane.packedCorrection.ptr, (const int3*) lane.items.ptr, numItems, (float4*) lane.buffers.posq,
                       mixed ? (float4*) lane.buffers.posq_correction : (float4*) nullptr);
    IXH_HIP_CHECK(hipGetLastError());
    check(ixh_state_clear(lane.state));
    hipLaunchKernelGGL(k_zero_doubles, dim3((s->numClusters+255)/256), dim3(256), 0, st, lane.energies.ptr, s->numClusters);
    const int flags = IXH_INCLUDE_FORCES | IXH_INCLUDE_ENERGY;
    if (lane.pair != nullptr)
        check(ixh_pair_execute_subset(lane.pair, &lane.buffers, s->pairTemplate.terms, flags, lane.energies.ptr, lane.hostSegments,
                                      numItems));
    if (lane.threebody != nullptr)
        check(ixh_threebody_execute_subset(lane.threebody, &lane.buffers, flags, lane.energies.ptr, lane.hostSegments, numItems));
    const int energyOffset = 3*(int) particles;
    hipLaunchKernelGGL(k_gather_results, dim3(4, numItems), dim3(256), 0, st, (const long long*) lane.buffers.force,
                       lane.buffers.force_stride, (const int3*) lane.items.ptr, (const int*) lane.segments.ptr, numItems,
                       (const double*) lane.energies.ptr, lane.results.ptr, energyOffset);
    IXH_HIP_CHECK(hipGetLastError());
    IXH_HIP_CHECK(hipMemcpyAsync(lane.hostResults, lane.results.ptr, sizeof(double)*(energyOffset+numItems), hipMemcpyDeviceToHost, st));
    IXH_HIP_CHECK(hipEventRecord(lane.done, st));
    lane.batch = batch;
    lane.busy = true;
    for (auto& r : batch)
        s->entries[r->entry]->inFlight = true;
    s->batches++;
    s->batchedRequests += numItems;
}

void finish_batch(ixh_server* s, Lane& lane) {
    const Clock::time_point now = Clock::now();
    int offset = 0;
    int particles = 0;
    for (auto& r : lane.batch)
        particles += s->entries[r->entry]->numParticles;
    const int energyOffset = 3*particles;
    for (size_t k = 0; k < lane.batch.size(); k++) {
        auto& r = lane.batch[k];
        Entry& e = *s->entries[r->entry];
        r->forces.assign(lane.hostResults+3*offset, lane.hostResults+3*(offset+e.numParticles));
        r->energy = lane.hostResults[energyOffset+k];
        r->completion = now;
        r->done = true;
        e.inFlight = false;
        offset += e.numParticles;
    }
    lane.batch.clear();
    lane.busy = false;
}

// Fixed-point forces of all particles and the energy of a periodic slot into one array.
__global__ void k_periodic_results(const long long* force, int stride, int n, const double* energy, double* out) {
    for (int k = blockIdx.x*blockDim.x+threadIdx.x; k < n; k += gridDim.x*blockDim.x)
        for (int c = 0; c < 3; c++)
            out[3*(size_t) k+c] = (double) force[(size_t) c*stride+k]/FIXED_POINT_SCALE;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        out[3*(size_t) n] = *energy;
}

// Evaluate a periodic request on a slot, asynchronously; the neighbour list is checked on the device.
void launch_periodic(ixh_server* s, PeriodicSlot& slot, const std::shared_ptr<Request>& r, bool deferCheck) {
    Entry& e = *s->entries[slot.entry];
    const int n = e.numParticles;
    const double* p = r->positions.data();
    for (int i = 0; i < n; i++) {
        float hx, hy, hz, lx, ly, lz;
        split(p[3*i], hx, lx);
        split(p[3*i+1], hy, ly);
        split(p[3*i+2], hz, lz);
        slot.hostPosq[i] = make_float4(hx, hy, hz, 0.0f);
        slot.hostCorrection[i] = make_float4(lx, ly, lz, 0.0f);
    }
    hipStream_t st = (hipStream_t) slot.buffers.stream;
    IXH_HIP_CHECK(hipMemcpyAsync((void*) slot.buffers.posq, slot.hostPosq, sizeof(float4)*n, hipMemcpyHostToDevice, st));
    if (s->precision == IXH_MIXED)
        IXH_HIP_CHECK(hipMemcpyAsync((void*) slot.buffers.posq_correction, slot.hostCorrection, sizeof(float4)*n, hipMemcpyHostToDevice, st));
    check(ixh_state_clear(slot.state));
    ixh_buffers b = slot.buffers;
    b.lambda = r->lambda;
    check(ixh_pair_execute(slot.pair, &b, 7, IXH_INCLUDE_FORCES | IXH_INCLUDE_ENERGY | (deferCheck ? IXH_DEFER_LIST_CHECK : 0)));
    hipLaunchKernelGGL(k_periodic_results, dim3(std::max(1, std::min(64, (n+255)/256))), dim3(256), 0, st,
                       (const long long*) b.force, b.force_stride, n, (const double*) b.energy, slot.results.ptr);
    IXH_HIP_CHECK(hipGetLastError());
    IXH_HIP_CHECK(hipMemcpyAsync(slot.hostResults, slot.results.ptr, sizeof(double)*(3*(size_t) n+1), hipMemcpyDeviceToHost, st));
    IXH_HIP_CHECK(hipEventRecord(slot.done, st));
    slot.request = r;
}

void finish_periodic(ixh_server* s, PeriodicSlot& slot) {
    auto r = slot.request;
    Entry& e = *s->entries[slot.entry];
    int valid = 1;
    check(ixh_pair_list_was_valid(slot.pair, &valid));
    if (!valid) {
        // the coordinates moved beyond the list's skin: evaluate again, rebuilding the list
        launch_periodic(s, slot, r, false);
        IXH_HIP_CHECK(hipEventSynchronize(slot.done));
    }
    const int n = e.numParticles;
    double lrc = 0;
    check(ixh_pair_long_range_energy(slot.pair, e.box, r->lambda, &lrc));
    r->forces.assign(slot.hostResults, slot.hostResults+3*(size_t) n);
    r->energy = slot.hostResults[3*(size_t) n]+lrc;
    r->completion = Clock::now();
    r->done = true;
    slot.request.reset();
}

void serve(ixh_server* s) {
    IXH_HIP_CHECK(hipSetDevice(s->device));
    std::deque<std::shared_ptr<Request>> pending;
    while (true) {
        {
            std::unique_lock<std::mutex> lock(s->mutex);
            bool anyBusy = s->lanes[0].busy || s->lanes[1].busy;
            for (auto& slot : s->periodicSlots)
                anyBusy = anyBusy || slot->request != nullptr;
            if (!s->stopping && s->incoming.empty() && pending.empty() && !anyBusy)
                s->submitted.wait(lock, [&] { return s->stopping || !s->incoming.empty(); });
            if (s->stopping && s->incoming.empty() && pending.empty() && !anyBusy)
                return;
            while (!s->incoming.empty()) {
                pending.push_back(s->incoming.front());
                s->incoming.pop_front();
            }
        }
        bool progress = false;
        // completions
        for (Lane& lane : s->lanes)
            if (lane.busy && hipEventQuery(lane.done) == hipSuccess) {
                std::lock_guard<std::mutex> lock(s->mutex);
                finish_batch(s, lane);
                progress = true;
            }
        for (auto& slot : s->periodicSlots)
            if (slot->request != nullptr && hipEventQuery(slot->done) == hipSuccess) {
                std::lock_guard<std::mutex> lock(s->mutex);
                try {
                    finish_periodic(s, *slot);
                }
                catch (const std::exception& ex) {
                    fail_request(s, slot->request, ex.what());
                    slot->request.reset();
                }
                progress = true;
            }
        if (progress)
            s->completed.notify_all();
        // admissions: every free lane takes the pending cluster requests that fit
        for (Lane& lane : s->lanes) {
            if (lane.busy)
                continue;
            std::vector<std::shared_ptr<Request>> batch;
            long long bytes = 0;
            for (auto it = pending.begin(); it != pending.end() && (int) batch.size() < s->maxBatch;) {
                Entry& e = *s->entries[(*it)->entry];
                bool taken = false;
                for (auto& q : batch)
                    taken = taken || q->entry == (*it)->entry;
                // (each lane has its own positions and forces, so a system may be in flight on both lanes at once)
                if (!e.periodic && !taken && (batch.empty() || bytes+e.footprint <= s->memoryBudget)) {
                    bytes += e.footprint;
                    batch.push_back(*it);
                    it = pending.erase(it);
                }
                else
                    ++it;
            }
            if (!batch.empty()) {
                try {
                    launch_batch(s, lane, batch);
                }
                catch (const std::exception& ex) {
                    std::lock_guard<std::mutex> lock(s->mutex);
                    for (auto& r : batch)
                        fail_request(s, r, ex.what());
                    s->completed.notify_all();
                }
                progress = true;
            }
        }
        // periodic requests: each takes a free slot of its system, in arrival order
        for (auto it = pending.begin(); it != pending.end();) {
            Entry& e = *s->entries[(*it)->entry];
            PeriodicSlot* free = nullptr;
            if (e.periodic)
                for (int k : e.slots)
                    if (s->periodicSlots[k]->request == nullptr) {
                        free = s->periodicSlots[k].get();
                        break;
                    }
            if (free == nullptr) {
                ++it;
                continue;
            }
            auto r = *it;
            it = pending.erase(it);
            try {
                launch_periodic(s, *free, r, true);
            }
            catch (const std::exception& ex) {
                std::lock_guard<std::mutex> lock(s->mutex);
                free->request.reset();
                fail_request(s, r, ex.what());
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

}  // namespace

extern "C" {

int ixh_server_create(int device, int precision, int max_batch, long long memory_budget, ixh_server** out) {
    return guarded([&] {
        if (out == nullptr || max_batch <= 0 || memory_budget <= 0)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_create: invalid argument");
        if (precision != IXH_SINGLE && precision != IXH_MIXED)
            throw Error(IXH_ERROR_UNSUPPORTED, "ixh_server_create: the server runs in single or mixed precision");
        ixh_server* s = new ixh_server();
        s->device = device;
        s->precision = precision;
        s->maxBatch = max_batch;
        s->memoryBudget = memory_budget;
        *out = s;
    });
}

int ixh_server_add_cluster(ixh_server* s, const ixh_pair_desc* pair, const ixh_threebody_desc* threebody, int* entry) {
    return guarded([&] {
        if (s == nullptr || pair == nullptr || entry == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_add_cluster: null argument");
        if (s->started)
            throw Error(IXH_ERROR_STATE, "ixh_server_add_cluster: the server has started");
        if (pair->method != IXH_NO_CUTOFF)
            throw Error(IXH_ERROR_UNSUPPORTED, "ixh_server_add_cluster: cluster systems use the NoCutoff method");
        bool first = true;
        for (auto& e : s->entries)
            first = first && e->periodic;
        if (first) {
            s->pairTemplate = *pair;
            s->haveThreebody = (threebody != nullptr);
            if (threebody != nullptr)
                s->threebodyTemplate = *threebody;
        }
        else if (!same_pair_settings(s->pairTemplate, *pair) || s->haveThreebody != (threebody != nullptr) ||
                 (threebody != nullptr && !same_threebody_settings(s->threebodyTemplate, *threebody)))
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
        e->footprint = cluster_footprint(*e);
        *entry = (int) s->entries.size();
        s->entries.push_back(std::move(e));
    });
}

int ixh_server_add_periodic(ixh_server* s, const ixh_pair_desc* pair, const double* box, int* entry) {
    return guarded([&] {
        if (s == nullptr || pair == nullptr || box == nullptr || entry == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_add_periodic: null argument");
        if (s->started)
            throw Error(IXH_ERROR_STATE, "ixh_server_add_periodic: the server has started");
        std::unique_ptr<Entry> e(new Entry());
        e->periodic = true;
        e->numParticles = pair->num_particles;
        std::copy(box, box+9, e->box);
        ixh_pair_desc d = *pair;
        d.precision = s->precision;
        d.default_box = box;
        const int index = (int) s->entries.size();
        const int n = e->numParticles;
        for (int k = 0; k < s->slotsPerPeriodic; k++) {
            std::unique_ptr<PeriodicSlot> slot(new PeriodicSlot());
            slot->entry = index;
            check(ixh_pair_create(s->device, &d, &slot->pair));
            check(ixh_state_create(s->device, n, s->precision, &slot->state));
            check(ixh_state_set_box(slot->state, box));
            check(ixh_state_get_buffers(slot->state, &slot->buffers));
            IXH_HIP_CHECK(hipHostMalloc((void**) &slot->hostPosq, sizeof(float4)*std::max(n, 1), hipHostMallocDefault));
            IXH_HIP_CHECK(hipHostMalloc((void**) &slot->hostCorrection, sizeof(float4)*std::max(n, 1), hipHostMallocDefault));
            IXH_HIP_CHECK(hipHostMalloc((void**) &slot->hostResults, sizeof(double)*(3*(size_t) n+1), hipHostMallocDefault));
            slot->results.reserve(3*(size_t) n+1);
            IXH_HIP_CHECK(hipEventCreateWithFlags(&slot->done, hipEventDisableTiming));
            e->slots.push_back((int) s->periodicSlots.size());
            s->periodicSlots.push_back(std::move(slot));
        }
        e->footprint = 200LL*e->numParticles;
        *entry = index;
        s->entries.push_back(std::move(e));
    });
}

int ixh_server_start(ixh_server* s) {
    return guarded([&] {
        if (s == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_start: null argument");
        if (s->started)
            return;
        DeviceGuard guard(s->device);
        build_catalog(s);
        // the first evaluation of every periodic slot builds its neighbour list (synchronously, before serving)
        for (auto& slot : s->periodicSlots) {
            ixh_buffers b = slot->buffers;
            check(ixh_pair_execute(slot->pair, &b, 7, IXH_INCLUDE_FORCES | IXH_INCLUDE_ENERGY));
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
                        fail_request(s, t.second, ex.what());
                s->completed.notify_all();
            }
        });
    });
}

int ixh_server_submit(ixh_server* s, int entry, const double* positions, double lambda, long long* ticket) {
    return guarded([&] {
        if (s == nullptr || positions == nullptr || ticket == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_submit: null argument");
        if (!s->started)
            throw Error(IXH_ERROR_STATE, "ixh_server_submit: the server has not started");
        if (entry < 0 || entry >= (int) s->entries.size())
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_submit: unknown system");
        auto r = std::make_shared<Request>();
        r->entry = entry;
        r->lambda = lambda;
        r->positions.assign(positions, positions+3*(size_t) s->entries[entry]->numParticles);
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
        *ticket = r->ticket;
    });
}

int ixh_server_wait(ixh_server* s, long long ticket, double* energy, double* forces, double* latency_ms) {
    return guarded([&] {
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
        if (energy != nullptr)
            *energy = r->energy;
        if (forces != nullptr)
            std::copy(r->forces.begin(), r->forces.end(), forces);
        if (latency_ms != nullptr)
            *latency_ms = std::chrono::duration<double, std::milli>(r->completion-r->arrival).count();
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

int ixh_server_replay(ixh_server* s, int n, const double* arrival, const int* entries, const double* lambdas, const long long* offsets,
                      const double* positions, double* latency_ms, double* energies, double* makespan_s) {
    return guarded([&] {
        if (s == nullptr || n < 0 || (n > 0 && (arrival == nullptr || entries == nullptr || offsets == nullptr || positions == nullptr)))
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_server_replay: null argument");
        std::vector<long long> tickets(n);
        const Clock::time_point start = Clock::now();
        // submit at the scheduled times (the latency clock of a request starts at its scheduled arrival)
        for (int k = 0; k < n; k++) {
            const Clock::time_point due = start+std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival[k]));
            while (Clock::now() < due) {
            }
            check(ixh_server_submit(s, entries[k], positions+offsets[k], lambdas != nullptr ? lambdas[k] : 1.0, &tickets[k]));
            std::lock_guard<std::mutex> lock(s->mutex);
            s->tickets[tickets[k]]->arrival = due;
        }
        Clock::time_point last = start;
        for (int k = 0; k < n; k++) {
            std::shared_ptr<Request> r;
            {
                std::lock_guard<std::mutex> lock(s->mutex);
                r = s->tickets[tickets[k]];
            }
            double e = 0, ms = 0;
            check(ixh_server_wait(s, tickets[k], &e, nullptr, &ms));
            if (latency_ms != nullptr)
                latency_ms[k] = ms;
            if (energies != nullptr)
                energies[k] = e;
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
            if (lane.pair != nullptr)
                ixh_pair_destroy(lane.pair);
            if (lane.threebody != nullptr)
                ixh_threebody_destroy(lane.threebody);
            if (lane.state != nullptr)
                ixh_state_destroy(lane.state);
            for (void* p : {(void*) lane.hostPacked, (void*) lane.hostPackedCorrection, (void*) lane.hostItems, (void*) lane.hostSegments,
                            (void*) lane.hostResults})
                if (p != nullptr)
                    (void) hipHostFree(p);
            if (lane.done != nullptr)
                (void) hipEventDestroy(lane.done);
        }
        for (auto& slot : s->periodicSlots) {
            if (slot->pair != nullptr)
                ixh_pair_destroy(slot->pair);
            if (slot->state != nullptr)
                ixh_state_destroy(slot->state);
            for (void* p : {(void*) slot->hostPosq, (void*) slot->hostCorrection, (void*) slot->hostResults})
                if (p != nullptr)
                    (void) hipHostFree(p);
            if (slot->done != nullptr)
                (void) hipEventDestroy(slot->done);
        }
        delete s;
    });
}

}  // extern "C"
