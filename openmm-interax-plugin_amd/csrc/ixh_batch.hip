// Batched evaluation of many independent systems (the batched configuration: uneven gas-phase clusters).
//
// No counterpart in the reference, whose kernels evaluate one System per Context.  The systems of a batch are
// concatenated into one segmented pair handle (NoCutoff tiles within each system only) and one segmented three-body
// handle (neighbour lists of static capacity), sharing one device state.  One evaluation is then a fixed set of
// kernels with no host synchronization, the pair half and the three-body half on two streams (forked and joined by
// events), captured once into a HIP graph and replayed; per-system energies are summed in a fixed order, so a
// system's energy and forces are the same alone and in any batch.  A tuner times the launch shapes of the pair tiles
// (workgroups of the persistent tile launch) and of the triple kernel (wavefronts and neighbour pairs per work item)
// and keeps the fastest.
#include "ixh_common.h"
#include "ixh_internal.h"

#include <algorithm>
#include <vector>

using namespace ixh;

namespace {

__global__ void k_fill_double(double* p, int n, double value) {
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n)
        p[i] = value;
}

// Zero the per-system energies of both halves of an evaluation and the three-body half's total.
__global__ void k_zero_energies(double* pair, double* three, int n, double* threeTotal) {
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n) {
        pair[i] = 0;
        three[i] = 0;
    }
    if (i == 0)
        *threeTotal = 0;
}

// After both halves: each system's energy is its pair energy plus its three-body energy, in that order (the sum a
// single stream forms), and the three-body total joins the state's energy.
__global__ void k_add_three_body(double* energies, const double* three, int n, double* total, const double* threeTotal) {
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n)
        energies[i] += three[i];
    if (i == 0)
        total[0] += *threeTotal;
}

void check(int status) {
    if (status != IXH_OK)
        throw Error(status, ixh_last_error());
}

}  // namespace

struct ixh_batch {
    int device = 0;
    int numSystems = 0;
    int numParticles = 0;
    int precision = IXH_MIXED;
    int pairTerms = 0;
    std::vector<int> particleOffsets;
    ixh_pair* pair = nullptr;
    ixh_threebody* threebody = nullptr;
    ixh_state* state = nullptr;
    ixh_buffers buffers;
    DeviceArray<double> energies;
    // The pair half of an evaluation runs on the state's stream and the three-body half beside it on `side` (forked
    // and joined by events, also inside the captured graph), with its own energies, added at the join.
    hipStream_t side = nullptr;
    hipEvent_t forked = nullptr, joined = nullptr, searched = nullptr;
    DeviceArray<double> threeEnergies, threeTotal;
    double* hostEnergies = nullptr;   // pinned
    int evaluated = 0;                // eager evaluations so far (the graph is captured after the first)
    int graphFlags = -1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    int useGraph = 1;
    int gridLimit = 4096;
};

namespace {

void release_side(ixh_batch* b) {
    if (b->side != nullptr)
        (void) hipStreamDestroy(b->side);
    if (b->forked != nullptr)
        (void) hipEventDestroy(b->forked);
    if (b->joined != nullptr)
        (void) hipEventDestroy(b->joined);
    if (b->searched != nullptr)
        (void) hipEventDestroy(b->searched);
    b->side = nullptr;
    b->forked = b->joined = b->searched = nullptr;
}

void destroy_graph(ixh_batch* b) {
    if (b->graphExec != nullptr)
        (void) hipGraphExecDestroy(b->graphExec);
    if (b->graph != nullptr)
        (void) hipGraphDestroy(b->graph);
    b->graphExec = nullptr;
    b->graph = nullptr;
    b->graphFlags = -1;
}

// The kernels of one evaluation, on the state's stream.
void enqueue(ixh_batch* b, int flags) {
    hipStream_t stream = (hipStream_t) b->buffers.stream;
    const dim3 grid((b->numSystems+255)/256);
    check(ixh_state_clear(b->state));
    if (b->pair == nullptr || b->threebody == nullptr) {
        hipLaunchKernelGGL(k_fill_double, grid, dim3(256), 0, stream, b->energies.ptr, b->numSystems, 0.0);
        IXH_HIP_CHECK(hipGetLastError());
        if (b->pair != nullptr)
            check(ixh_pair_execute_segments(b->pair, &b->buffers, b->pairTerms, flags, b->energies.ptr));
        if (b->threebody != nullptr)
            check(ixh_threebody_execute_segments(b->threebody, &b->buffers, flags, b->energies.ptr));
        return;
    }
    hipLaunchKernelGGL(k_zero_energies, grid, dim3(256), 0, stream, b->energies.ptr, b->threeEnergies.ptr, b->numSystems,
                       b->threeTotal.ptr);
    IXH_HIP_CHECK(hipGetLastError());
    IXH_HIP_CHECK(hipEventRecord(b->forked, stream));
    IXH_HIP_CHECK(hipStreamWaitEvent(b->side, b->forked, 0));
    ixh_buffers three = b->buffers;
    three.stream = b->side;
    three.energy = b->threeTotal.ptr;
    check(ixh_threebody_execute_segments(b->threebody, &three, flags, b->threeEnergies.ptr));
    // the pair half starts once the three-body half has its neighbour lists, so that the (latency-bound) search runs
    // alone and the pair tiles share the device with the triple kernel only
    IXH_HIP_CHECK(hipStreamWaitEvent(stream, b->searched, 0));
    check(ixh_pair_execute_segments(b->pair, &b->buffers, b->pairTerms, flags, b->energies.ptr));
    IXH_HIP_CHECK(hipEventRecord(b->joined, b->side));
    IXH_HIP_CHECK(hipStreamWaitEvent(stream, b->joined, 0));
    hipLaunchKernelGGL(k_add_three_body, grid, dim3(256), 0, stream, b->energies.ptr, (const double*) b->threeEnergies.ptr,
                       b->numSystems, b->buffers.energy, (const double*) b->threeTotal.ptr);
    IXH_HIP_CHECK(hipGetLastError());
}

void run(ixh_batch* b, int flags) {
    hipStream_t stream = (hipStream_t) b->buffers.stream;
    // the first evaluation builds the lists and sizes every buffer; later ones replay a captured graph (profiling
    // records events around kernels, so it runs eagerly)
    if (!b->useGraph || b->evaluated == 0 || profiling_enabled()) {
        enqueue(b, flags);
        b->evaluated++;
        return;
    }
    if (b->graphExec == nullptr || b->graphFlags != flags) {
        destroy_graph(b);
        IXH_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            enqueue(b, flags);
        }
        catch (...) {
            hipGraph_t discard = nullptr;
            (void) hipStreamEndCapture(stream, &discard);
            if (discard != nullptr)
                (void) hipGraphDestroy(discard);
            throw;
        }
        IXH_HIP_CHECK(hipStreamEndCapture(stream, &b->graph));
        IXH_HIP_CHECK(hipGraphInstantiate(&b->graphExec, b->graph, nullptr, nullptr, 0));
        b->graphFlags = flags;
    }
    IXH_HIP_CHECK(hipGraphLaunch(b->graphExec, stream));
    b->evaluated++;
}

}  // namespace

extern "C" {

int ixh_batch_create(int device, const ixh_batch_desc* d, ixh_batch** out) {
    return guarded([&] {
        if (d == nullptr || out == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_create: null argument");
        if (d->num_systems <= 0 || d->particle_offsets == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_create: a batch needs at least one system");
        if (d->pair == nullptr && d->threebody == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_create: no forces");
        if (d->threebody != nullptr && d->site_offsets == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_create: the three-body force needs site offsets");
        DeviceGuard guard(device);
        ixh_batch* b = new ixh_batch();
        try {
            b->device = device;
            b->numSystems = d->num_systems;
            b->particleOffsets.assign(d->particle_offsets, d->particle_offsets+d->num_systems+1);
            b->numParticles = b->particleOffsets.back();
            b->precision = d->precision;
            if (d->pair != nullptr) {
                if (d->pair->num_particles != b->numParticles)
                    throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_create: the pair forces cover a different number of particles");
                check(ixh_pair_create_segmented(device, d->pair, d->num_systems, d->particle_offsets, &b->pair));
                b->pairTerms = d->pair->terms;
            }
            if (d->threebody != nullptr) {
                if (d->threebody->num_particles != b->numParticles)
                    throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_create: the three-body force covers a different number of particles");
                check(ixh_threebody_create_segmented(device, d->threebody, d->num_systems, d->site_offsets, &b->threebody));
            }
            check(ixh_state_create(device, b->numParticles, d->precision, &b->state));
            check(ixh_state_get_buffers(b->state, &b->buffers));
            b->energies.reserve(b->numSystems);
            if (b->pair != nullptr && b->threebody != nullptr) {
                b->threeEnergies.reserve(b->numSystems);
                b->threeTotal.reserve(1);
                IXH_HIP_CHECK(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
                IXH_HIP_CHECK(hipEventCreateWithFlags(&b->forked, hipEventDisableTiming));
                IXH_HIP_CHECK(hipEventCreateWithFlags(&b->joined, hipEventDisableTiming));
                IXH_HIP_CHECK(hipEventCreateWithFlags(&b->searched, hipEventDisableTiming));
                threebody_set_search_event(b->threebody, b->searched);
            }
            IXH_HIP_CHECK(hipHostMalloc((void**) &b->hostEnergies, sizeof(double)*b->numSystems, hipHostMallocDefault));
        }
        catch (...) {
            if (b->pair != nullptr)
                ixh_pair_destroy(b->pair);
            if (b->threebody != nullptr)
                ixh_threebody_destroy(b->threebody);
            if (b->state != nullptr)
                ixh_state_destroy(b->state);
            if (b->hostEnergies != nullptr)
                (void) hipHostFree(b->hostEnergies);
            release_side(b);
            delete b;
            throw;
        }
        *out = b;
    });
}

int ixh_batch_destroy(ixh_batch* b) {
    return guarded([&] {
        if (b == nullptr)
            return;
        DeviceGuard guard(b->device);
        (void) hipDeviceSynchronize();
        destroy_graph(b);
        if (b->pair != nullptr)
            ixh_pair_destroy(b->pair);
        if (b->threebody != nullptr)
            ixh_threebody_destroy(b->threebody);
        if (b->state != nullptr)
            ixh_state_destroy(b->state);
        if (b->hostEnergies != nullptr)
            (void) hipHostFree(b->hostEnergies);
        release_side(b);
        delete b;
    });
}

int ixh_batch_set_positions(ixh_batch* b, const double* positions) {
    return guarded([&] {
        if (b == nullptr || positions == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_set_positions: null argument");
        DeviceGuard guard(b->device);
        check(ixh_state_set_positions(b->state, positions));
    });
}

int ixh_batch_get_buffers(ixh_batch* b, ixh_buffers* out) {
    return guarded([&] {
        if (b == nullptr || out == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_get_buffers: null argument");
        *out = b->buffers;
    });
}

int ixh_batch_execute(ixh_batch* b, int flags) {
    return guarded([&] {
        if (b == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_execute: null argument");
        DeviceGuard guard(b->device);
        run(b, flags);
    });
}

int ixh_batch_get_energies(ixh_batch* b, double* energies) {
    return guarded([&] {
        if (b == nullptr || energies == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_get_energies: null argument");
        DeviceGuard guard(b->device);
        hipStream_t stream = (hipStream_t) b->buffers.stream;
        IXH_HIP_CHECK(hipMemcpyAsync(b->hostEnergies, b->energies.ptr, sizeof(double)*b->numSystems, hipMemcpyDeviceToHost, stream));
        IXH_HIP_CHECK(hipStreamSynchronize(stream));
        std::copy(b->hostEnergies, b->hostEnergies+b->numSystems, energies);
    });
}

int ixh_batch_get_forces(ixh_batch* b, double* forces) {
    return guarded([&] {
        if (b == nullptr || forces == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_get_forces: null argument");
        DeviceGuard guard(b->device);
        check(ixh_state_get_forces(b->state, forces));
    });
}

int ixh_batch_set_graph(ixh_batch* b, int enable) {
    return guarded([&] {
        if (b == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_set_graph: null argument");
        DeviceGuard guard(b->device);
        IXH_HIP_CHECK(hipStreamSynchronize((hipStream_t) b->buffers.stream));
        destroy_graph(b);
        b->useGraph = (enable != 0);
    });
}

int ixh_batch_kernel_time(ixh_batch* b, double* pair_ms, long long* pair_launches, double* triple_ms, long long* triple_launches) {
    return guarded([&] {
        if (b == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_kernel_time: null argument");
        double ms = 0;
        long long launches = 0;
        if (b->pair != nullptr)
            check(ixh_pair_kernel_time(b->pair, &ms, &launches));
        if (pair_ms != nullptr)
            *pair_ms = ms;
        if (pair_launches != nullptr)
            *pair_launches = launches;
        ms = 0;
        launches = 0;
        if (b->threebody != nullptr)
            check(ixh_threebody_kernel_time(b->threebody, &ms, &launches));
        if (triple_ms != nullptr)
            *triple_ms = ms;
        if (triple_launches != nullptr)
            *triple_launches = launches;
    });
}

int ixh_batch_reset_kernel_time(ixh_batch* b) {
    return guarded([&] {
        if (b == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_reset_kernel_time: null argument");
        if (b->pair != nullptr)
            check(ixh_pair_reset_kernel_time(b->pair));
        if (b->threebody != nullptr)
            check(ixh_threebody_reset_kernel_time(b->threebody));
    });
}

int ixh_batch_get_shape(ixh_batch* b, ixh_batch_shape* shape) {
    return guarded([&] {
        if (b == nullptr || shape == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_get_shape: null argument");
        shape->pair_workgroups = b->gridLimit;
        shape->triple_waves = 0;
        shape->triple_pairs_per_item = 0;
        if (b->threebody != nullptr)
            threebody_launch_shape(b->threebody, &shape->triple_waves, &shape->triple_pairs_per_item);
    });
}

int ixh_batch_set_shape(ixh_batch* b, const ixh_batch_shape* shape) {
    return guarded([&] {
        if (b == nullptr || shape == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_set_shape: null argument");
        if (shape->pair_workgroups < 1)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_set_shape: the pair launch needs at least one workgroup");
        DeviceGuard guard(b->device);
        IXH_HIP_CHECK(hipStreamSynchronize((hipStream_t) b->buffers.stream));
        b->gridLimit = shape->pair_workgroups;
        if (b->pair != nullptr)
            pair_set_grid_limit(b->pair, b->gridLimit);
        if (b->threebody != nullptr)
            threebody_set_launch_shape(b->threebody, shape->triple_waves, shape->triple_pairs_per_item);
        destroy_graph(b);
    });
}

int ixh_batch_autotune(ixh_batch* b, int flags, ixh_batch_shape* chosen) {
    return guarded([&] {
        if (b == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_batch_autotune: null argument");
        DeviceGuard guard(b->device);
        hipStream_t stream = (hipStream_t) b->buffers.stream;
        if (b->evaluated == 0)
            run(b, flags);   // builds the tile lists
        hipEvent_t start, stop;
        IXH_HIP_CHECK(hipEventCreate(&start));
        IXH_HIP_CHECK(hipEventCreate(&stop));
        // milliseconds per evaluation of the current shape: one eager evaluation sizes its buffers, the next captures
        // its graph, five replays are timed
        auto time_shape = [&]() {
            destroy_graph(b);
            enqueue(b, flags);
            run(b, flags);
            IXH_HIP_CHECK(hipStreamSynchronize(stream));
            IXH_HIP_CHECK(hipEventRecord(start, stream));
            for (int rep = 0; rep < 5; rep++)
                run(b, flags);
            IXH_HIP_CHECK(hipEventRecord(stop, stream));
            IXH_HIP_CHECK(hipEventSynchronize(stop));
            float ms = 0;
            IXH_HIP_CHECK(hipEventElapsedTime(&ms, start, stop));
            return ms/5;
        };
        try {
            if (b->pair != nullptr) {
                const int tiles = pair_num_tiles(b->pair);
                const int candidates[] = {256, 512, 1024, 2048, 4096, 8192, 16384};
                float best = 1e30f;
                int bestLimit = b->gridLimit;
                for (int limit : candidates) {
                    if (limit > 2*std::max(1, (tiles+3)/4) && limit != candidates[0])
                        break;   // more workgroups than tiles: the same launch as the previous candidate
                    pair_set_grid_limit(b->pair, limit);
                    const float ms = time_shape();
                    if (ms < best) {
                        best = ms;
                        bestLimit = limit;
                    }
                }
                b->gridLimit = bestLimit;
                pair_set_grid_limit(b->pair, bestLimit);
            }
            if (b->threebody != nullptr) {
                int bestWaves = 4, bestChunk = 16384;
                threebody_launch_shape(b->threebody, &bestWaves, &bestChunk);
                float best = 1e30f;
                for (int waves : {4, 8, 2, 1})
                    for (int chunk : {8192, 16384, 32768, 65536}) {
                        IXH_HIP_CHECK(hipStreamSynchronize(stream));
                        threebody_set_launch_shape(b->threebody, waves, chunk);
                        const float ms = time_shape();
                        if (ms < best) {
                            best = ms;
                            bestWaves = waves;
                            bestChunk = chunk;
                        }
                    }
                IXH_HIP_CHECK(hipStreamSynchronize(stream));
                threebody_set_launch_shape(b->threebody, bestWaves, bestChunk);
            }
        }
        catch (...) {
            (void) hipEventDestroy(start);
            (void) hipEventDestroy(stop);
            throw;
        }
        (void) hipEventDestroy(start);
        (void) hipEventDestroy(stop);
        destroy_graph(b);
        if (chosen != nullptr) {
            chosen->pair_workgroups = b->gridLimit;
            chosen->triple_waves = 0;
            chosen->triple_pairs_per_item = 0;
            if (b->threebody != nullptr)
                threebody_launch_shape(b->threebody, &chosen->triple_waves, &chosen->triple_pairs_per_item);
        }
    });
}

}  // extern "C"
