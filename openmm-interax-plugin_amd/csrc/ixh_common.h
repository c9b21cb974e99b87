// Shared host and device helpers of the interax_hip library (gfx950).
#ifndef IXH_COMMON_H
#define IXH_COMMON_H

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "interax_hip.h"

namespace ixh {

// ------------------------------------------------------------------------------------------------------------
// Errors: every C entry point catches ixh::Error and stores its message for ixh_last_error().
// ------------------------------------------------------------------------------------------------------------

struct Error : std::runtime_error {
    int code;
    Error(int code, const std::string& message) : std::runtime_error(message), code(code) {}
};

void set_last_error(const std::string& message);

#define IXH_HIP_CHECK(call)                                                                                   \
    do {                                                                                                      \
        hipError_t ixh_status_ = (call);                                                                      \
        if (ixh_status_ != hipSuccess)                                                                        \
            throw ::ixh::Error(IXH_ERROR_HIP, std::string(#call) + ": " + hipGetErrorString(ixh_status_));    \
    } while (0)

// Run a body, translating exceptions into a status code.
template <class F>
int guarded(F&& body) {
    try {
        body();
        return IXH_OK;
    }
    catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    }
    catch (const std::exception& e) {
        set_last_error(e.what());
        return IXH_ERROR_HIP;
    }
    catch (...) {
        set_last_error("unknown error");
        return IXH_ERROR_HIP;
    }
}

// Bytes of device memory held by the library's DeviceArrays (all handles, all devices), for the server's admission.
extern std::atomic<long long> deviceBytes;

// Device memory owned by a handle.
template <class T>
struct DeviceArray {
    T* ptr = nullptr;
    size_t count = 0;
    size_t bytes = 0;
    DeviceArray() = default;
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    ~DeviceArray() { release(); }
    void release() {
        if (ptr != nullptr) {
            (void) hipFree(ptr);
            deviceBytes.fetch_sub((long long) bytes, std::memory_order_relaxed);
        }
        ptr = nullptr;
        count = 0;
        bytes = 0;
    }
    // Grow to at least n elements (contents are not kept).
    void reserve(size_t n) {
        if (n <= count && ptr != nullptr)
            return;
        release();
        const size_t size = (n == 0 ? 1 : n)*sizeof(T);
        IXH_HIP_CHECK(hipMalloc((void**) &ptr, size));
        count = n;
        bytes = size;
        deviceBytes.fetch_add((long long) size, std::memory_order_relaxed);
    }
    void upload(const T* host, size_t n, hipStream_t stream) {
        reserve(n);
        if (n > 0)
            IXH_HIP_CHECK(hipMemcpyAsync(ptr, host, n*sizeof(T), hipMemcpyHostToDevice, stream));
    }
};

// Pinned host staging of the small host-to-device copies of an asynchronous call, in a ring of buffers: a call waits
// only until the copy made from the same buffer RING calls earlier has been read, not for the call just before it
// (whose copy is queued behind the kernels before it).  Growing one buffer grows all of them.
struct StagingRing {
    static constexpr int RING = 8;
    int* buffer[RING] = {};
    hipEvent_t copied[RING] = {};
    size_t capacity = 0;
    int next = 0;
    StagingRing() = default;
    StagingRing(const StagingRing&) = delete;
    StagingRing& operator=(const StagingRing&) = delete;
    ~StagingRing() { release(); }
    void release() {
        for (int k = 0; k < RING; k++) {
            if (copied[k] != nullptr) {
                (void) hipEventSynchronize(copied[k]);
                (void) hipEventDestroy(copied[k]);
            }
            if (buffer[k] != nullptr)
                (void) hipHostFree(buffer[k]);
            buffer[k] = nullptr;
            copied[k] = nullptr;
        }
        capacity = 0;
    }
    // A buffer of at least n ints that no pending copy reads; `slot` identifies it for recorded().
    int* acquire(size_t n, int& slot) {
        if (n > capacity) {
            for (int k = 0; k < RING; k++) {
                if (copied[k] != nullptr)
                    IXH_HIP_CHECK(hipEventSynchronize(copied[k]));
                if (buffer[k] != nullptr)
                    IXH_HIP_CHECK(hipHostFree(buffer[k]));
                buffer[k] = nullptr;
            }
            capacity = std::max<size_t>(n, 1024);
            for (int k = 0; k < RING; k++)
                IXH_HIP_CHECK(hipHostMalloc((void**) &buffer[k], sizeof(int)*capacity, hipHostMallocDefault));
        }
        slot = next;
        next = (next+1)%RING;
        if (copied[slot] == nullptr)
            IXH_HIP_CHECK(hipEventCreateWithFlags(&copied[slot], hipEventDisableTiming));
        else
            IXH_HIP_CHECK(hipEventSynchronize(copied[slot]));
        return buffer[slot];
    }
    // The copies from buffer `slot` have been queued on `stream`.
    void recorded(int slot, hipStream_t stream) { IXH_HIP_CHECK(hipEventRecord(copied[slot], stream)); }
};

// Sets the device for the lifetime of the object and restores the previous one.
struct DeviceGuard {
    int previous = 0;
    explicit DeviceGuard(int device) {
        IXH_HIP_CHECK(hipGetDevice(&previous));
        if (previous != device)
            IXH_HIP_CHECK(hipSetDevice(device));
    }
    ~DeviceGuard() { (void) hipSetDevice(previous); }
};

// Whether launches of the main kernels are bracketed by events (ixh_set_profiling).
bool profiling_enabled();

// HIP events around the launches of one kernel, summed on request.
struct KernelTimer {
    std::vector<hipEvent_t> pool;   // start, stop, start, stop, ...
    size_t used = 0;
    double accumulated = 0;
    long long launches = 0;
    hipEvent_t next() {
        if (used == pool.size()) {
            hipEvent_t e;
            IXH_HIP_CHECK(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[used++];
    }
    // Record a start event; returns false when profiling is off.
    bool start(hipStream_t stream) {
        if (!profiling_enabled())
            return false;
        if (used+2 > 4096)
            collect();
        IXH_HIP_CHECK(hipEventRecord(next(), stream));
        return true;
    }
    void stop(hipStream_t stream) {
        IXH_HIP_CHECK(hipEventRecord(next(), stream));
        launches++;
    }
    void collect() {
        for (size_t k = 0; k+1 < used; k += 2) {
            IXH_HIP_CHECK(hipEventSynchronize(pool[k+1]));
            float ms = 0;
            IXH_HIP_CHECK(hipEventElapsedTime(&ms, pool[k], pool[k+1]));
            accumulated += ms;
        }
        used = 0;
    }
    void reset() {
        collect();
        accumulated = 0;
        launches = 0;
    }
    ~KernelTimer() {
        for (hipEvent_t e : pool)
            (void) hipEventDestroy(e);
    }
};

constexpr double FIXED_POINT_SCALE = 4294967296.0;   // 2^32 units per kJ/mol/nm, as OpenMM's GPU platforms

// ------------------------------------------------------------------------------------------------------------
// Device helpers
// ------------------------------------------------------------------------------------------------------------

// Wave64 lane rotation: lane l receives the value of lane (l+1) mod 64.  DPP wave_rol:1 is a GFX9-family control
// (CDNA keeps the wave-wide shifts and rotations that RDNA dropped).
__device__ __forceinline__ float wave_rotate_down(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x134, 0xF, 0xF, false));
}
__device__ __forceinline__ int wave_rotate_down(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x134, 0xF, 0xF, false);
}
__device__ __forceinline__ unsigned wave_rotate_down(unsigned v) {
    return (unsigned) __builtin_amdgcn_update_dpp(0, (int) v, 0x134, 0xF, 0xF, false);
}
__device__ __forceinline__ double wave_rotate_down(double v) {
    long long bits = __builtin_bit_cast(long long, v);
    int lo = (int) (bits & 0xffffffffLL), hi = (int) (bits >> 32);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x134, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x134, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (long long) (((unsigned long long) (unsigned) hi << 32) | (unsigned) lo));
}

// Convert a force component to 64-bit fixed point.
__device__ __forceinline__ long long to_fixed(double v) {
    return (long long) (v*FIXED_POINT_SCALE);
}
// FP32, five instructions: the high word is floor(v) (v_cvt_flr_i32_f32, one instruction where floor and convert
// take two), the low word the fraction v - floor(v), taken in FP64 (v_fract_f64), times 2^32.  The FP64 fraction of
// a 24-bit v is exact unless -2^-29 < v < 0, where it is 1+v rounded to FP64 (kept below 1): the result is then
// floor(v 2^32) to within one unit.  Valid for |v| < 2^31.  (The FP32 fraction, one instruction less, would lose up to
// 2^8 units for -1/2 < v < 0: 3e-8 kJ/mol/nm, more than the forces of distant pairs can afford.)
__device__ __forceinline__ long long to_fixed(float v) {
    int hi;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(hi) : "v"(v));
    const unsigned lo = (unsigned) (__builtin_amdgcn_fract((double) v)*FIXED_POINT_SCALE);
    return (long long) (((unsigned long long) (unsigned) hi << 32) | lo);
}

// The lanes of the wavefront for which p holds (the compare's lane mask itself, with no copy through a VGPR).
__device__ __forceinline__ unsigned long long wave_ballot(bool p) {
    return __builtin_amdgcn_ballot_w64(p);
}
// base plus the number of lanes below this one whose bit is set in mask (v_mbcnt_lo and v_mbcnt_hi).
__device__ __forceinline__ unsigned lanes_below(unsigned long long mask, unsigned base) {
    return __builtin_amdgcn_mbcnt_hi((unsigned) (mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned) mask, base));
}

__device__ __forceinline__ void add_fixed(long long* p, long long v) {
    if (v != 0)
        atomicAdd((unsigned long long*) p, (unsigned long long) v);
}

// Fast FP32 intrinsics: v_rcp_f32, v_rsq_f32, v_exp_f32 (2^x).
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ double fast_rcp(double x) { return 1.0/x; }
__device__ __forceinline__ double fast_rsqrt(double x) { return 1.0/sqrt(x); }
__device__ __forceinline__ double fast_exp2(double x) { return exp2(x); }

// Wave-wide sum of a double, result valid in every lane.
__device__ __forceinline__ double wave_sum(double v) {
    for (int offset = 32; offset > 0; offset >>= 1)
        v += __shfl_xor(v, offset, 64);
    return v;
}
// FP32 sum over the 64 lanes, every lane receiving the same order-fixed total (all lanes active): DPP within each
// row of 16 lanes (quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror, row_mirror; each step adds two partial sums
// that are equal up to the order of one commutative addition), then the four row sums through two lane exchanges.
__device__ __forceinline__ float wave_sum(float v) {
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, false));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, false));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x141, 0xF, 0xF, false));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x140, 0xF, 0xF, false));
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

}  // namespace ixh

#endif
