// Device state: the per-Context buffers a platform hands to execute() (positions as posq, a fixed-point force
// buffer, an FP64 energy accumulator, the box and the alchemical parameter), for hosts that do not own such
// buffers themselves.  Positions are laid out as OpenMM's GPU platforms lay them out: float4 posq in single
// precision, float4 posq plus a float4 correction in mixed precision, double4 in double precision.
#include "ixh_common.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <vector>

namespace ixh {

static thread_local std::string lastError;
// set by any thread, read by the server's thread among others
static std::atomic<bool> profiling{false};
std::atomic<long long> deviceBytes{0};

bool profiling_enabled() {
    return profiling.load(std::memory_order_relaxed);
}

void set_last_error(const std::string& message) {
    lastError = message;
}

}  // namespace ixh

using namespace ixh;

namespace {

__global__ void k_clear(long long* force, long long n, double* energy) {
    long long i = (long long) blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n)
        force[i] = 0;
    if (i == 0)
        energy[0] = 0;
}

}  // namespace

struct ixh_state {
    int device = 0;
    int numParticles = 0;
    int precision = IXH_MIXED;
    int stride = 0;
    hipStream_t stream = nullptr;      // the stream of every operation: the state's own, or one set by the caller
    hipStream_t ownStream = nullptr;
    DeviceArray<char> posq;
    DeviceArray<float4> correction;
    DeviceArray<long long> force;
    DeviceArray<double> energy;
    double box[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double lambda = 1.0;
    std::vector<char> hostPosq;
    std::vector<float4> hostCorrection;
    std::vector<long long> hostForce;
};

extern "C" {

const char* ixh_last_error(void) {
    return ixh::lastError.c_str();
}

int ixh_version(int* major, int* minor) {
    return guarded([&] {
        if (major != nullptr)
            *major = IXH_VERSION_MAJOR;
        if (minor != nullptr)
            *minor = IXH_VERSION_MINOR;
    });
}

int ixh_set_profiling(int enable) {
    return guarded([&] { ixh::profiling.store(enable != 0, std::memory_order_relaxed); });
}

int ixh_device_count(int* count) {
    return guarded([&] {
        if (count == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_device_count: null pointer");
        int n = 0;
        IXH_HIP_CHECK(hipGetDeviceCount(&n));
        *count = n;
    });
}

int ixh_state_create(int device, int numParticles, int precision, ixh_state** out) {
    return guarded([&] {
        if (out == nullptr || numParticles < 0)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_state_create: invalid argument");
        if (precision != IXH_SINGLE && precision != IXH_MIXED && precision != IXH_DOUBLE)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_state_create: illegal precision");
        DeviceGuard guard(device);
        ixh_state* s = new ixh_state();
        try {
            s->device = device;
            s->numParticles = numParticles;
            s->precision = precision;
            s->stride = ((numParticles+63)/64)*64;
            IXH_HIP_CHECK(hipStreamCreateWithFlags(&s->ownStream, hipStreamNonBlocking));
            s->stream = s->ownStream;
            size_t posBytes = (precision == IXH_DOUBLE ? sizeof(double4) : sizeof(float4));
            s->posq.reserve(posBytes*(size_t) (numParticles > 0 ? numParticles : 1));
            IXH_HIP_CHECK(hipMemset(s->posq.ptr, 0, s->posq.count));
            if (precision == IXH_MIXED) {
                s->correction.reserve(numParticles);
                IXH_HIP_CHECK(hipMemset(s->correction.ptr, 0, sizeof(float4)*s->correction.count));
            }
            s->force.reserve(3*(size_t) s->stride);
            s->energy.reserve(1);
            IXH_HIP_CHECK(hipMemset(s->force.ptr, 0, sizeof(long long)*s->force.count));
            IXH_HIP_CHECK(hipMemset(s->energy.ptr, 0, sizeof(double)));
        }
        catch (...) {
            if (s->ownStream != nullptr)
                (void) hipStreamDestroy(s->ownStream);
            delete s;
            throw;
        }
        *out = s;
    });
}

int ixh_state_destroy(ixh_state* s) {
    return guarded([&] {
        if (s == nullptr)
            return;
        DeviceGuard guard(s->device);
        if (s->stream != s->ownStream)
            (void) hipStreamSynchronize(s->stream);
        if (s->ownStream != nullptr) {
            (void) hipStreamSynchronize(s->ownStream);
            (void) hipStreamDestroy(s->ownStream);
        }
        delete s;
    });
}

int ixh_state_set_positions(ixh_state* s, const double* positions) {
    return guarded([&] {
        if (s == nullptr || (positions == nullptr && s->numParticles > 0))
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_state_set_positions: invalid argument");
        DeviceGuard guard(s->device);
        IXH_HIP_CHECK(hipStreamSynchronize(s->stream));
        int n = s->numParticles;
        if (s->precision == IXH_DOUBLE) {
            s->hostPosq.resize(sizeof(double4)*(size_t) n);
            double4* p = (double4*) s->hostPosq.data();
            for (int i = 0; i < n; i++)
                p[i] = make_double4(positions[3*i], positions[3*i+1], positions[3*i+2], 0.0);
        }
        else {
            s->hostPosq.resize(sizeof(float4)*(size_t) n);
            float4* p = (float4*) s->hostPosq.data();
            if (s->precision == IXH_MIXED)
                s->hostCorrection.resize(n);
            for (int i = 0; i < n; i++) {
                float x = (float) positions[3*i], y = (float) positions[3*i+1], z = (float) positions[3*i+2];
                p[i] = make_float4(x, y, z, 0.0f);
                if (s->precision == IXH_MIXED)
                    s->hostCorrection[i] = make_float4((float) (positions[3*i]-(double) x), (float) (positions[3*i+1]-(double) y),
                                                       (float) (positions[3*i+2]-(double) z), 0.0f);
            }
        }
        if (n > 0) {
            IXH_HIP_CHECK(hipMemcpyAsync(s->posq.ptr, s->hostPosq.data(), s->hostPosq.size(), hipMemcpyHostToDevice, s->stream));
            if (s->precision == IXH_MIXED)
                IXH_HIP_CHECK(hipMemcpyAsync(s->correction.ptr, s->hostCorrection.data(), sizeof(float4)*(size_t) n,
                                             hipMemcpyHostToDevice, s->stream));
        }
        IXH_HIP_CHECK(hipStreamSynchronize(s->stream));
    });
}

int ixh_state_set_box(ixh_state* s, const double* box) {
    return guarded([&] {
        if (s == nullptr || box == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_state_set_box: invalid argument");
        std::memcpy(s->box, box, sizeof(s->box));
    });
}

int ixh_state_set_lambda(ixh_state* s, double lambda) {
    return guarded([&] {
        if (s == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_state_set_lambda: invalid argument");
        s->lambda = lambda;
    });
}

int ixh_state_set_stream(ixh_state* s, void* stream) {
    return guarded([&] {
        if (s == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_state_set_stream: invalid argument");
        DeviceGuard guard(s->device);
        IXH_HIP_CHECK(hipStreamSynchronize(s->stream));
        s->stream = (stream != nullptr ? (hipStream_t) stream : s->ownStream);
    });
}

int ixh_state_clear(ixh_state* s) {
    return guarded([&] {
        if (s == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_state_clear: invalid argument");
        DeviceGuard guard(s->device);
        long long n = 3LL*s->stride;
        hipLaunchKernelGGL(k_clear, dim3((unsigned) std::max(1LL, (n+255)/256)), dim3(256), 0, s->stream, s->force.ptr, n, s->energy.ptr);
        IXH_HIP_CHECK(hipGetLastError());
    });
}

int ixh_state_synchronize(ixh_state* s) {
    return guarded([&] {
        if (s == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_state_synchronize: invalid argument");
        DeviceGuard guard(s->device);
        IXH_HIP_CHECK(hipStreamSynchronize(s->stream));
    });
}

int ixh_state_get_forces(ixh_state* s, double* forces) {
    return guarded([&] {
        if (s == nullptr || (forces == nullptr && s->numParticles > 0))
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_state_get_forces: invalid argument");
        DeviceGuard guard(s->device);
        s->hostForce.resize(3*(size_t) s->stride);
        IXH_HIP_CHECK(hipMemcpyAsync(s->hostForce.data(), s->force.ptr, sizeof(long long)*3*(size_t) s->stride,
                                     hipMemcpyDeviceToHost, s->stream));
        IXH_HIP_CHECK(hipStreamSynchronize(s->stream));
        for (int i = 0; i < s->numParticles; i++)
            for (int k = 0; k < 3; k++)
                forces[3*i+k] = (double) s->hostForce[(size_t) k*s->stride+i]/FIXED_POINT_SCALE;
    });
}

int ixh_state_get_energy(ixh_state* s, double* energy) {
    return guarded([&] {
        if (s == nullptr || energy == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_state_get_energy: invalid argument");
        DeviceGuard guard(s->device);
        IXH_HIP_CHECK(hipMemcpyAsync(energy, s->energy.ptr, sizeof(double), hipMemcpyDeviceToHost, s->stream));
        IXH_HIP_CHECK(hipStreamSynchronize(s->stream));
    });
}

int ixh_state_get_buffers(ixh_state* s, ixh_buffers* out) {
    return guarded([&] {
        if (s == nullptr || out == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_state_get_buffers: invalid argument");
        out->num_particles = s->numParticles;
        out->precision = s->precision;
        out->posq = s->posq.ptr;
        out->posq_correction = (s->precision == IXH_MIXED ? s->correction.ptr : nullptr);
        out->force = s->force.ptr;
        out->force_stride = s->stride;
        out->energy = s->energy.ptr;
        std::memcpy(out->box, s->box, sizeof(s->box));
        out->lambda = s->lambda;
        out->stream = s->stream;
        out->atom_index = nullptr;
    });
}

}  // extern "C"
