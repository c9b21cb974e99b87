// The fused InteraxFF pair forces on gfx950: neighbour list, candidate masks, the pair kernels and the dispersion
// long range correction.
//
// Reference behaviour (the contract): ReferenceInteraxPairIxn.h:53-99 (pairs, cutoff, soft core, switching,
// alchemical scaling), ReferenceInteraxKernels.cpp:41-198 (the three force kernels), InteraxPairForce.cpp:113-146
// (validation), InteraxDispersionForceImpl.cpp:86-149 (long range correction).
//
// Layout in HBM.  Particles are sorted along a Morton curve of their fractional coordinates into blocks of 64 (one
// wavefront); per sorted slot a kernel reads one 16-byte position in the frame of its block and three 16-byte
// parameter vectors.  A tile pairs the 64 particles of a block I with 64 partner slots: the block itself (the
// diagonal tile) or a chunk of the partner list of I.  With a cutoff, the partner list of I holds the particles of
// blocks J > I that lie within cutoff + skin of the bounding box of I (a particle-box test, which keeps about twice
// as many useful pairs per tile as a box-box test); without a cutoff it is every block J > I.  Each tile has a
// 64 x 64 candidate mask, built with the list: bit (i, j) is set when the pair is not excluded, both slots hold
// particles, and (with a cutoff) the pair was within cutoff + skin.
//
// Kernels.  Without a cutoff every candidate interacts: one wavefront evaluates a tile in 64 rotation steps (lane l
// pairs i = l with j = (l+s) % 64, j read from LDS, the j force accumulators following their particle around the
// wave with a DPP rotation).  With a cutoff each lane walks the set bits of its own mask row, tests the distance and
// evaluates only the pairs inside the cutoff, so lanes do useful work instead of idling on distant pairs; the forces
// on j are added in LDS as 64-bit fixed point.
//
// Determinism.  Forces are summed per tile in FP32 in a fixed order, converted to 64-bit fixed point and added with
// integer atomics; each tile writes its FP64 energy to its own slot and the slots are summed in tile order.  Results
// therefore do not depend on the launch shape, on which wave ran which tile, or on other systems sharing a launch.
#include "ixh_pair_math.h"
#include "ixh_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include <rocprim/rocprim.hpp>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>
#include <vector>

using namespace ixh;

namespace {

constexpr int BLOCK = 64;
// Per-slot flags: the original particle index, the alchemical flag in bit 30, -1 for a padding slot.
constexpr int INDEX_MASK = 0x3fffffff;
constexpr int ALCHEMICAL_BIT = 0x40000000;
constexpr double SKIN = 0.1;   // neighbour list padding (nm); rebuilt when a particle has moved skin/2
// Tile records are (block I, partner-chunk offset or -1 for the diagonal tile); bit 30 of the first word marks a tile
// whose pairs need their own minimum image (the box is too small for one image shift per block pair).
constexpr int TILE_PAIR_IMAGE = 1 << 30;
constexpr int TILE_BLOCK_MASK = TILE_PAIR_IMAGE-1;

template <typename Real>
struct Vec4T;
template <>
struct Vec4T<float> {
    typedef float4 type;
};
template <>
struct Vec4T<double> {
    typedef double4 type;
};

// ------------------------------------------------------------------------------------------------------------
// Geometry helpers
// ------------------------------------------------------------------------------------------------------------

struct BoxD {
    double ax, bx, by, cx, cy, cz;   // reduced form: a = (ax,0,0), b = (bx,by,0), c = (cx,cy,cz)
};

// Minimum image of d, reduced along c, b, a in turn, with every operation rounded as in
// ReferenceForce::getDeltaRPeriodic (no fused multiply-add), so that the result is bit-identical to the reference.
__device__ __forceinline__ void reference_min_image(double& dx, double& dy, double& dz, const BoxD& b) {
    double s2 = floor(__dadd_rn(__ddiv_rn(dz, b.cz), 0.5));
    dx = __dsub_rn(dx, __dmul_rn(b.cx, s2));
    dy = __dsub_rn(dy, __dmul_rn(b.cy, s2));
    dz = __dsub_rn(dz, __dmul_rn(b.cz, s2));
    double s1 = floor(__dadd_rn(__ddiv_rn(dy, b.by), 0.5));
    dx = __dsub_rn(dx, __dmul_rn(b.bx, s1));
    dy = __dsub_rn(dy, __dmul_rn(b.by, s1));
    double s0 = floor(__dadd_rn(__ddiv_rn(dx, b.ax), 0.5));
    dx = __dsub_rn(dx, __dmul_rn(b.ax, s0));
}

template <typename Real>
__device__ __forceinline__ void min_image(Real& dx, Real& dy, Real& dz, Real ax, Real bx, Real by, Real cx, Real cy, Real cz,
                                          Real iax, Real iby, Real icz) {
    Real s2 = floor(dz*icz+Real(0.5));
    dx -= s2*cx;
    dy -= s2*cy;
    dz -= s2*cz;
    Real s1 = floor(dy*iby+Real(0.5));
    dx -= s1*bx;
    dy -= s1*by;
    Real s0 = floor(dx*iax+Real(0.5));
    dx -= s0*ax;
}

// FP64 position of particle i from the platform's posq layout.
__device__ __forceinline__ void load_position(const void* posq, const float4* correction, int precision, int i,
                                              double& x, double& y, double& z) {
    if (precision == IXH_DOUBLE) {
        double4 p = ((const double4*) posq)[i];
        x = p.x;
        y = p.y;
        z = p.z;
    }
    else {
        float4 p = ((const float4*) posq)[i];
        x = p.x;
        y = p.y;
        z = p.z;
        if (correction != nullptr) {
            float4 c = correction[i];
            x += c.x;
            y += c.y;
            z += c.z;
        }
    }
}

// ------------------------------------------------------------------------------------------------------------
// Neighbour list kernels (run when the list is rebuilt)
// ------------------------------------------------------------------------------------------------------------

__device__ __forceinline__ unsigned spread_bits(unsigned v) {
    v &= 0x3ff;
    v = (v | (v << 16)) & 0x030000FF;
    v = (v | (v << 8)) & 0x0300F00F;
    v = (v | (v << 4)) & 0x030C30C3;
    v = (v | (v << 2)) & 0x09249249;
    return v;
}

struct KeyArgs {
    const void* posq;
    const float4* correction;
    int precision;
    int n;
    int periodic;
    BoxD box;
    double lo[3], scale[3];   // nonperiodic: fractional coordinate = (x - lo)*scale
};

__global__ void k_morton_keys(KeyArgs a, unsigned* keys, int* values, double* pos3) {
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= a.n)
        return;
    double x, y, z;
    load_position(a.posq, a.correction, a.precision, i, x, y, z);
    pos3[3*i] = x;
    pos3[3*i+1] = y;
    pos3[3*i+2] = z;
    double f[3];
    if (a.periodic) {
        double fz = z/a.box.cz;
        double fy = (y-fz*a.box.cy)/a.box.by;
        double fx = (x-fz*a.box.cx-fy*a.box.bx)/a.box.ax;
        f[0] = fx-floor(fx);
        f[1] = fy-floor(fy);
        f[2] = fz-floor(fz);
    }
    else {
        f[0] = (x-a.lo[0])*a.scale[0];
        f[1] = (y-a.lo[1])*a.scale[1];
        f[2] = (z-a.lo[2])*a.scale[2];
    }
    unsigned c[3];
    for (int k = 0; k < 3; k++) {
        int v = (int) (f[k]*1024.0);
        c[k] = (unsigned) min(max(v, 0), 1023);
    }
    keys[i] = spread_bits(c[0]) | (spread_bits(c[1]) << 1) | (spread_bits(c[2]) << 2);
    values[i] = i;
}

__global__ void k_bounding_box(const void* posq, const float4* correction, int precision, int n, double* out) {
    // one workgroup of 256 threads: min and max of the coordinates
    __shared__ double sMin[3][256], sMax[3][256];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double p[3];
        load_position(posq, correction, precision, i, p[0], p[1], p[2]);
        for (int k = 0; k < 3; k++) {
            lo[k] = fmin(lo[k], p[k]);
            hi[k] = fmax(hi[k], p[k]);
        }
    }
    for (int k = 0; k < 3; k++) {
        sMin[k][threadIdx.x] = lo[k];
        sMax[k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride)
            for (int k = 0; k < 3; k++) {
                sMin[k][threadIdx.x] = fmin(sMin[k][threadIdx.x], sMin[k][threadIdx.x+stride]);
                sMax[k][threadIdx.x] = fmax(sMax[k][threadIdx.x], sMax[k][threadIdx.x+stride]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) {
            out[k] = sMin[k][0];
            out[3+k] = sMax[k][0];
        }
}

// Slot bookkeeping after the sort: padding slots get -1; slotOf is the inverse permutation.
__global__ void k_finish_perm(int* perm, int n, int padded, int* slotOf) {
    int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= padded)
        return;
    if (s >= n)
        perm[s] = -1;
    else
        slotOf[perm[s]] = s;
}

template <typename Real>
__global__ void k_gather_params(const int* perm, int padded, const typename Vec4T<Real>::type* pA, const typename Vec4T<Real>::type* pB,
                                const typename Vec4T<Real>::type* pC, const int* flags, typename Vec4T<Real>::type* sA,
                                typename Vec4T<Real>::type* sB, typename Vec4T<Real>::type* sC, int* sFlags) {
    int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= padded)
        return;
    int i = perm[s];
    typedef typename Vec4T<Real>::type R4;
    if (i < 0) {
        R4 zero = {0, 0, 0, 0};
        sA[s] = zero;
        sB[s] = zero;
        sC[s] = zero;
        sFlags[s] = -1;
        return;
    }
    sA[s] = pA[i];
    sB[s] = pB[i];
    sC[s] = pC[i];
    sFlags[s] = flags[i];
}

__global__ void k_gather_int(const int* perm, int padded, const int* src, int* dst) {
    int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= padded)
        return;
    int i = perm[s];
    dst[s] = (i < 0 ? 0 : src[i]);
}

// Bounding box of each block, in the frame of its first particle (periodic images taken about it).
struct BlockBoundsArgs {
    const double* pos3;
    const int* perm;
    int numBlocks;
    int periodic;
    BoxD box;
};

__global__ void k_block_bounds(BlockBoundsArgs a, double* bounds /* 6 per block: center, half extent */) {
    int b = blockIdx.x*blockDim.x+threadIdx.x;
    if (b >= a.numBlocks)
        return;
    int first = a.perm[b*BLOCK];
    double o[3] = {a.pos3[3*first], a.pos3[3*first+1], a.pos3[3*first+2]};
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int k = 0; k < BLOCK; k++) {
        int i = a.perm[b*BLOCK+k];
        if (i < 0)
            break;
        double d[3] = {a.pos3[3*i]-o[0], a.pos3[3*i+1]-o[1], a.pos3[3*i+2]-o[2]};
        if (a.periodic)
            reference_min_image(d[0], d[1], d[2], a.box);
        for (int c = 0; c < 3; c++) {
            lo[c] = fmin(lo[c], d[c]);
            hi[c] = fmax(hi[c], d[c]);
        }
    }
    for (int c = 0; c < 3; c++) {
        bounds[6*b+c] = o[c]+0.5*(lo[c]+hi[c]);
        bounds[6*b+3+c] = 0.5*(hi[c]-lo[c]);
    }
}

// Smallest distance between the boxes of blocks I and J, over the periodic images within one lattice vector of the
// reduced centre difference (a conservative test: it never drops a pair that can be within range).
__device__ double block_distance2(const double* bounds, int I, int J, int periodic, const BoxD& box) {
    double d[3] = {bounds[6*J]-bounds[6*I], bounds[6*J+1]-bounds[6*I+1], bounds[6*J+2]-bounds[6*I+2]};
    double h[3] = {bounds[6*I+3]+bounds[6*J+3], bounds[6*I+4]+bounds[6*J+4], bounds[6*I+5]+bounds[6*J+5]};
    if (!periodic) {
        double s = 0;
        for (int c = 0; c < 3; c++) {
            double g = fmax(0.0, fabs(d[c])-h[c]);
            s += g*g;
        }
        return s;
    }
    reference_min_image(d[0], d[1], d[2], box);
    double best = 1e300;
    for (int i = -1; i <= 1; i++)
        for (int j = -1; j <= 1; j++)
            for (int k = -1; k <= 1; k++) {
                double e[3] = {d[0]+i*box.ax+j*box.bx+k*box.cx, d[1]+j*box.by+k*box.cy, d[2]+k*box.cz};
                double s = 0;
                for (int c = 0; c < 3; c++) {
                    double g = fmax(0.0, fabs(e[c])-h[c]);
                    s += g*g;
                }
                best = fmin(best, s);
            }
    return best;
}

// Smallest distance between a point d (relative to the centre of block I) and the box of I, over the periodic images
// within one lattice vector of the reduced displacement.
__device__ double point_box_distance2(double dx, double dy, double dz, const double* h, int periodic, const BoxD& box) {
    if (!periodic) {
        double gx = fmax(0.0, fabs(dx)-h[0]), gy = fmax(0.0, fabs(dy)-h[1]), gz = fmax(0.0, fabs(dz)-h[2]);
        return gx*gx+gy*gy+gz*gz;
    }
    reference_min_image(dx, dy, dz, box);
    double best = 1e300;
    for (int i = -1; i <= 1; i++)
        for (int j = -1; j <= 1; j++)
            for (int k = -1; k <= 1; k++) {
                double ex = dx+i*box.ax+j*box.bx+k*box.cx, ey = dy+j*box.by+k*box.cy, ez = dz+k*box.cz;
                double gx = fmax(0.0, fabs(ex)-h[0]), gy = fmax(0.0, fabs(ey)-h[1]), gz = fmax(0.0, fabs(ez)-h[2]);
                best = fmin(best, gx*gx+gy*gy+gz*gz);
            }
    return best;
}

// Count (pass 0) or write (pass 1) the candidate blocks J > I of each block I: box distance below the list cutoff.
__global__ void k_find_blocks(const double* bounds, int numBlocks, int periodic, BoxD box, double listCutoff2, int* rowCount,
                              const int* rowStart, int* blocks) {
    int I = blockIdx.x;
    int lane = threadIdx.x;   // 64 threads
    int written = 0;
    for (int J0 = I+1; J0 < numBlocks; J0 += 64) {
        int J = J0+lane;
        bool hit = (J < numBlocks) && block_distance2(bounds, I, J, periodic, box) < listCutoff2;
        unsigned long long ballot = __ballot(hit);
        if (blocks != nullptr && hit)
            blocks[rowStart[I]+written+__popcll(ballot & ((1ULL << lane)-1))] = J;
        written += __popcll(ballot);
    }
    if (blocks == nullptr && lane == 0)
        rowCount[I] = written;
}

// Count (pass 0) or write (pass 1) the partners of each block I: the particles of its candidate blocks within the
// list cutoff of the bounding box of I, in ascending slot order.
__global__ void k_find_partners(const double* bounds, const double* pos3, const int* perm, const int* blockStart,
                                const int* blocks, int periodic, BoxD box, double listCutoff2, int* count,
                                const int* partnerStart, int* partners) {
    int I = blockIdx.x;
    int lane = threadIdx.x;   // 64 threads
    const double c[3] = {bounds[6*I], bounds[6*I+1], bounds[6*I+2]};
    const double h[3] = {bounds[6*I+3], bounds[6*I+4], bounds[6*I+5]};
    int written = 0;
    for (int k = blockStart[I]; k < blockStart[I+1]; k++) {
        int slot = blocks[k]*BLOCK+lane;
        int atom = perm[slot];
        bool hit = false;
        if (atom >= 0)
            hit = point_box_distance2(pos3[3*atom]-c[0], pos3[3*atom+1]-c[1], pos3[3*atom+2]-c[2], h, periodic, box) < listCutoff2;
        unsigned long long ballot = __ballot(hit);
        if (partners != nullptr && hit)
            partners[partnerStart[I]+written+__popcll(ballot & ((1ULL << lane)-1))] = slot;
        written += __popcll(ballot);
    }
    if (partners == nullptr && lane == 0)
        count[I] = written;
}

__device__ __forceinline__ bool excluded(const int* offsets, const int* partners, int i, int j) {
    int lo = offsets[i], hi = offsets[i+1];
    while (lo < hi) {
        int mid = (lo+hi) >> 1;
        int v = partners[mid];
        if (v == j)
            return true;
        if (v < j)
            lo = mid+1;
        else
            hi = mid;
    }
    return false;
}

// The candidate mask of every tile (one wavefront per tile, lane = i): bit j is set when the pair is not excluded,
// both slots hold particles, i != j, and (with a cutoff) the pair is within the list cutoff.  Masks of diagonal tiles
// are symmetric.
struct MaskArgs {
    const int2* tiles;
    const int* partners;     // slot lists; tile.y < 0 marks the diagonal tile
    const int* perm;
    const double* pos3;
    const int* exclusionOffsets;
    const int* exclusionPartners;
    int useCutoff;
    int periodic;
    BoxD box;
    double listCutoff2;
    const double* bounds;    // per block: centre, half extent (periodic cutoff lists)
    double imageLimit;       // half the smallest box width minus the cutoff and twice the skin
};

// Diameter of the bounding box of block b.
__device__ __forceinline__ double block_extent(const double* bounds, int b) {
    const double* h = bounds+6*b+3;
    return 2*sqrt(h[0]*h[0]+h[1]*h[1]+h[2]*h[2]);
}

// The candidate masks and partner slots of every tile; with periodic bounds, also whether the tile's pairs need
// their own minimum image (tileImage[t] = 1): one image shift per block pair is exact for every pair in range when
// cutoff + (extent_I + skin) + (extent_J + skin) < half the box width (particles move at most skin/2 before a rebuild).
__global__ void k_build_masks(MaskArgs a, unsigned long long* masks, int* tileSlots, int* tileImage) {
    __shared__ int sAtom[BLOCK];
    __shared__ double sPos[BLOCK][3];
    int t = blockIdx.x;
    int lane = threadIdx.x;
    int2 tile = a.tiles[t];
    int jslot = (tile.y < 0 ? tile.x*BLOCK+lane : a.partners[tile.y+lane]);
    int ja = (jslot < 0 ? -1 : a.perm[jslot]);
    tileSlots[(size_t) t*BLOCK+lane] = jslot;
    if (a.useCutoff && a.periodic) {
        double extent = (jslot < 0 ? 0.0 : block_extent(a.bounds, jslot/BLOCK));
        for (int offset = 32; offset > 0; offset >>= 1)
            extent = fmax(extent, __shfl_xor(extent, offset, 64));
        if (lane == 0)
            tileImage[t] = (block_extent(a.bounds, tile.x)+extent >= a.imageLimit) ? 1 : 0;
    }
    else if (lane == 0)
        tileImage[t] = 0;
    sAtom[lane] = ja;
    if (ja >= 0) {
        sPos[lane][0] = a.pos3[3*ja];
        sPos[lane][1] = a.pos3[3*ja+1];
        sPos[lane][2] = a.pos3[3*ja+2];
    }
    __syncthreads();
    int ia = a.perm[tile.x*BLOCK+lane];
    unsigned long long m = 0;
    if (ia >= 0) {
        double x = a.pos3[3*ia], y = a.pos3[3*ia+1], z = a.pos3[3*ia+2];
        for (int b = 0; b < BLOCK; b++) {
            int other = sAtom[b];
            if (other < 0 || other == ia)
                continue;
            if (excluded(a.exclusionOffsets, a.exclusionPartners, ia, other))
                continue;
            if (a.useCutoff) {
                double dx = sPos[b][0]-x, dy = sPos[b][1]-y, dz = sPos[b][2]-z;
                if (a.periodic)
                    reference_min_image(dx, dy, dz, a.box);
                if (dx*dx+dy*dy+dz*dz >= a.listCutoff2)
                    continue;
            }
            m |= 1ULL << b;
        }
    }
    masks[(size_t) t*BLOCK+lane] = m;
}

// Candidate list of a tile in rotation order: for s = 0..63 the pairs (i, (i+s) % 64) whose mask bit is set, so that
// consecutive entries have distinct i and distinct j.  Within the diagonal tile s runs over 1..32 (s = 32 for
// i < 32 only), which meets every unordered pair once.  Entries are i << 6 | j; each tile's list is padded with
// 0xffff to a multiple of 4 entries.  With offsets == nullptr only the counts (before padding) are written.
// Rows of the per-tile arrays (masks, partner slots) in a new tile order: row t of the output is row order[t].
__global__ void k_gather_tile_rows(const int* order, const unsigned long long* masks, const int* slots, unsigned long long* masksOut,
                                   int* slotsOut) {
    const size_t from = (size_t) order[blockIdx.x]*BLOCK+threadIdx.x, to = (size_t) blockIdx.x*BLOCK+threadIdx.x;
    masksOut[to] = masks[from];
    slotsOut[to] = slots[from];
}

// The candidate pairs of a tile in rotation order (pairs (i, i+s mod 64) for s = 1, 2, ...), padded to a multiple
// of 4 with 0xffff.  The tile kernel reads a list 256 entries per step, 4 consecutive u16 per lane, and takes them
// in 4 sub-steps; each chunk of 256 (or the last, shorter one, of 4m entries) is stored transposed so that sub-step
// s of lane l gets entry s m + l: in a sub-step consecutive lanes hold consecutive entries of the rotation order,
// whose particles lie in distinct LDS banks.  Without offsets only the counts are written.
__global__ void k_candidate_lists(const int2* tiles, const unsigned long long* masks, int* counts, const int* offsets,
                                  unsigned short* lists) {
    const int t = blockIdx.x;
    const int lane = threadIdx.x;
    const bool diag = (tiles[t].y < 0);
    const unsigned long long m = masks[(size_t) t*BLOCK+lane];
    const int sBegin = (diag ? 1 : 0), sEnd = (diag ? 33 : 64);
    int total = 0;
    for (int s = sBegin; s < sEnd; s++) {
        const int j = (lane+s) & 63;
        const bool in = ((m >> j) & 1ULL) != 0 && !(diag && s == 32 && lane >= 32);
        total += __popcll(__ballot(in));
    }
    if (offsets == nullptr) {
        if (lane == 0)
            counts[t] = total;
        return;
    }
    const int start = offsets[t];
    const int padded = (total+3) & ~3;
    // entry k of the list -> its position
    auto position = [padded](int k) {
        const int chunk = k & ~255, kk = k & 255;
        const int rows = min(256, padded-chunk)/4;   // lanes that read this chunk
        return chunk+(kk % rows)*4+kk/rows;
    };
    int count = 0;
    for (int s = sBegin; s < sEnd; s++) {
        const int j = (lane+s) & 63;
        const bool in = ((m >> j) & 1ULL) != 0 && !(diag && s == 32 && lane >= 32);
        const unsigned long long ballot = __ballot(in);
        if (in)
            lists[start+position(count+__popcll(ballot & ((1ULL << lane)-1)))] = (unsigned short) ((lane << 6) | j);
        count += __popcll(ballot);
    }
    if (lane < padded-total)
        lists[start+position(total+lane)] = 0xffff;
}

// ------------------------------------------------------------------------------------------------------------
// Per-evaluation kernels
// ------------------------------------------------------------------------------------------------------------

// The System index of a slot's particle (the platform's atom index array, or the identity) rides in the fourth
// component of its block-frame position, refreshed with the positions on every call: the InteraxOrder2 rule needs it,
// and a platform may reorder its particles between calls.  In FP32 the integer's bits are carried as they are.
__device__ __forceinline__ float encode_index(int i, float) { return __int_as_float(i); }
__device__ __forceinline__ double encode_index(int i, double) { return (double) i; }
__device__ __forceinline__ int decode_index(float w) { return __float_as_int(w); }
__device__ __forceinline__ int decode_index(double w) { return (int) w; }

// The neighbour-list check of a call, made in the same pass as its block frames (CHECK): flag is set if a particle
// moved more than sqrt(limit2) since the last build, and the flag of the other frame is cleared for the next call.
struct ListCheck {
    const double* buildPos;   // particle order
    double limit2;
    int* flag;
    int* otherFlag;
};

// Block frames: the centre of block b is the position of its first particle; every particle is stored relative to
// the centre of its block (nearest image), in the arithmetic type.  Every particle has exactly one slot, so with
// CHECK every particle's displacement is tested once.
template <typename Real, bool CHECK>
__global__ void k_local_positions(const void* posq, const float4* correction, int precision, const int* perm, int padded,
                                  int periodic, BoxD box, double* centers, typename Vec4T<Real>::type* local, ListCheck check,
                                  const int* atomIndex) {
    int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (CHECK && s == 0)
        *check.otherFlag = 0;
    if (s >= padded)
        return;
    int b = s/BLOCK;
    int first = perm[b*BLOCK];
    double cx, cy, cz;
    load_position(posq, correction, precision, first, cx, cy, cz);
    if (s%BLOCK == 0) {
        centers[3*b] = cx;
        centers[3*b+1] = cy;
        centers[3*b+2] = cz;
    }
    int i = perm[s];
    typedef typename Vec4T<Real>::type R4;
    if (i < 0) {
        R4 far = {Real(0), Real(0), Real(0), Real(0)};
        local[s] = far;
        return;
    }
    double x, y, z;
    load_position(posq, correction, precision, i, x, y, z);
    if (CHECK) {
        double ex = x-check.buildPos[3*i], ey = y-check.buildPos[3*i+1], ez = z-check.buildPos[3*i+2];
        if (periodic)
            reference_min_image(ex, ey, ez, box);
        if (ex*ex+ey*ey+ez*ez > check.limit2)
            *check.flag = 1;
    }
    double dx = x-cx, dy = y-cy, dz = z-cz;
    if (periodic)
        reference_min_image(dx, dy, dz, box);
    R4 v = {Real(dx), Real(dy), Real(dz), encode_index(atomIndex != nullptr ? atomIndex[i] : i, Real(0))};
    local[s] = v;
}

template <typename Real>
struct TileArgs {
    typedef typename Vec4T<Real>::type R4;
    const int2* tiles;       // (block I, offset of the partner chunk, or -1 for the diagonal tile)
    int numTiles;
    const int* partners;
    const int* tileSlots;    // the partner slot of every lane of every tile (-1 for padding)
    const unsigned long long* masks;
    const int* listOffsets;  // candidate list of tile t: lists[listOffsets[t] .. listOffsets[t+1]), padded to 4 entries
    const unsigned short* lists;
    const int* perm;         // slot -> particle (the platform's current order)
    const R4* local;
    const double* centers;
    const R4* pA;    // m1, m2, m3, alpha
    const R4* pB;    // dispersion width, exchange amplitude, exchange width, 1/exchange width
    const R4* pC;    // k, l, short-range width, 1/short-range width
    const int* flags;   // particle index (current order) | alchemical bit, -1 for padding; in LDS the System index
    // type tables (FP32 only): pA then holds wd, amp, K/2, L/2 and cls the class indices of each slot
    const int* cls;
    const float* tables;
    int tableFloats;
    int numDisp, numExch, numSr, exchOffset, srOffset;
    int highOrder;   // some overlap row of the tables is of degree 8
    int singleBranch;   // every width pair of the tables takes one overlap branch only
    int ordered;     // the overlap of a pair depends on which particle has the lower System index (InteraxOrder2)
    float beta2;
    int n;
    // geometry
    int method;
    int periodic;
    BoxD box;
    Real boxR[9];   // ax, bx, by, cx, cy, cz, 1/ax, 1/by, 1/cz
    double boxInv[3];   // 1/ax, 1/by, 1/cz
    Real cutoff2;
    double cutoff;
    Real band;      // |r^2 - rc^2| < band*rc^2 is decided exactly in FP64
    int useSwitch;
    Real switchDistance, invSwitchWidth;
    int useSoftcore;
    Real softcoreTerm;   // (1 - lambda) R^6
    Real lambda;
    PairSettings settings;
    // the platform buffers
    const void* posq;
    const float4* correction;
    int precision;
    long long* force;
    int stride;
    double* waveEnergy;   // per wave (all-pairs tiles) or per workgroup (cutoff tiles): energy and pair count, in tile order
    int* waveCount;
    double* tileEnergy;   // segmented handles: the energy of every tile, for the per-segment sums
    const int* tileIndex;  // subsets of segments: the tiles to evaluate (numTiles of them); nullptr: all tiles
    const int2* items;     // cutoff lists: the tile range [x, y) of every wave; nullptr: ranges split evenly
};

// Copy the type tables into LDS (all threads of the workgroup) and describe them.
template <typename Real>
__device__ __forceinline__ PairTables load_tables(const TileArgs<Real>& a, float* lds) {
    for (int i = threadIdx.x; i < a.tableFloats; i += blockDim.x)
        lds[i] = a.tables[i];
    __syncthreads();
    PairTables tb;
    tb.disp = (const float4*) lds;
    tb.exch = (const OverlapRow*) (lds+a.exchOffset);
    tb.sr = (const OverlapRow*) (lds+a.srOffset);
    tb.numDisp = a.numDisp;
    tb.numExch = a.numExch;
    tb.numSr = a.numSr;
    tb.beta2 = a.beta2;
    tb.shared = a.settings.shared_overlap;
    tb.highOrder = a.highOrder;
    tb.singleBranch = a.singleBranch;
    return tb;
}

// The particles of one side (i or j) of a tile in LDS.  With type tables only positions, the packed per-particle
// values and the class indices are kept.
template <typename Real, bool TABLES>
struct SideLds {
    typedef typename Vec4T<Real>::type R4;
    // Positions in separate x, y, z (and w) arrays: lanes reading 64 different particles of a side then touch 64
    // consecutive 4-byte words (no bank conflicts), where 16-byte records would put four particles in each bank.
    Real px[BLOCK], py[BLOCK], pz[BLOCK];
    float pw[BLOCK];                // with type tables the class indices, as bits
    Real wd[TABLES ? BLOCK : 1];    // with type tables: the dispersion width
    R4 A[TABLES ? 1 : BLOCK];
    R4 B[TABLES ? 1 : BLOCK];
    R4 C[TABLES ? 1 : BLOCK];
    int flags[BLOCK];
};

template <typename Real, bool TABLES>
__device__ __forceinline__ void store_pos(SideLds<Real, TABLES>& s, int l, const typename Vec4T<Real>::type& v) {
    s.px[l] = v.x;
    s.py[l] = v.y;
    s.pz[l] = v.z;
    s.pw[l] = (float) v.w;
}
template <typename Real, bool TABLES>
__device__ __forceinline__ typename Vec4T<Real>::type load_pos(const SideLds<Real, TABLES>& s, int l) {
    typename Vec4T<Real>::type v = {s.px[l], s.py[l], s.pz[l], Real(0)};
    return v;
}

template <typename Real, bool TABLES>
struct ParticleType {
    typedef PairParticle<Real> type;
};
template <>
struct ParticleType<float, true> {
    typedef TableParticle type;
};

// LDS flags of a slot: its System index (orig) with the alchemical bit of its flags, or -1 for padding.
__device__ __forceinline__ int lds_flags(int flags, int orig) {
    return (flags < 0 ? -1 : (flags & ALCHEMICAL_BIT) | orig);
}

// The fixed table form (k_pair_queue with FIXED): in place of its class word, a particle in LDS carries the byte
// offsets, from the start of the tables, of its part of a pair's rows -- the dispersion offset in the low 16 bits, the
// overlap offset in the high 16 -- in the role it has in the tile (i or j), so that the offsets of a pair's rows are
// the sum of the two particles' words.  The rows of pair (i, j) are dispersion class pair (d_i, d_j) and overlap
// class pair (x_j, x_i), the orders pair_energy_table uses when the index order does not matter.
template <int TERMS>
__device__ __forceinline__ int fixed_role_word(const TileArgs<float>& a, int cls, bool iRole) {
    constexpr bool EXCH = (TERMS & IXH_TERM_EXCHANGE) != 0;
    constexpr int PAIR_BYTES = 4*OVERLAP_PAIR_FLOATS;
    const int d = cls & 0xff;
    const int x = (EXCH ? cls >> 8 : cls >> 16) & 0xff;
    if (iRole)
        return d*a.numDisp*16 + ((4*(EXCH ? a.exchOffset : a.srOffset)+x*PAIR_BYTES) << 16);
    return d*16 + ((x*(EXCH ? a.numExch : a.numSr)*PAIR_BYTES) << 16);
}

// Stage the parameters of a slot whose position is already in LDS; orig is its System index.  Returns the slot's
// flags (particle index in the current order, for the force buffer).  FIXED: the i-side role word of the fixed
// table form in place of the class word.
template <typename Real, bool TABLES, int TERMS = 0, bool FIXED = false>
__device__ __forceinline__ int stage_particle(SideLds<Real, TABLES>& s, int lane, const TileArgs<Real>& a, int slot, int orig) {
    if constexpr (TABLES) {
        s.wd[lane] = a.pA[slot].x;
        if constexpr (FIXED)
            s.pw[lane] = __int_as_float(fixed_role_word<TERMS>(a, a.cls[slot], true));
        else
            s.pw[lane] = __int_as_float(a.cls[slot]);
    }
    else {
        s.A[lane] = a.pA[slot];
        s.B[lane] = a.pB[slot];
        s.C[lane] = a.pC[slot];
    }
    const int flags = a.flags[slot];
    s.flags[lane] = lds_flags(flags, orig);
    return flags;
}

__device__ __forceinline__ void fetch_particle(const SideLds<float, true>& s, int l, TableParticle& P) {
    P.wd = s.wd[l];
    P.cls = __float_as_int(s.pw[l]);
}
template <typename Real>
__device__ __forceinline__ void fetch_particle(const SideLds<Real, false>& s, int l, PairParticle<Real>& P) {
    const typename Vec4T<Real>::type A = s.A[l], B = s.B[l], C = s.C[l];
    P.m1 = A.x; P.m2 = A.y; P.m3 = A.z; P.alpha = A.w;
    P.wd = B.x; P.amp = B.y; P.wx = B.z; P.iwx = B.w;
    P.k = C.x; P.l = C.y; P.ws = C.z; P.iws = C.w;
}

// Exact membership test of the reference (ReferenceInteraxPairIxn.h:64-70) in FP64.
__device__ __noinline__ bool exact_within_cutoff(const void* posq, const float4* correction, int precision, int periodic,
                                                 BoxD box, double cutoff, int oi, int oj) {
    double xi, yi, zi, xj, yj, zj;
    load_position(posq, correction, precision, oi, xi, yi, zi);
    load_position(posq, correction, precision, oj, xj, yj, zj);
    double dx = __dsub_rn(xj, xi), dy = __dsub_rn(yj, yi), dz = __dsub_rn(zj, zi);
    if (periodic)
        reference_min_image(dx, dy, dz, box);
    double r2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
    return __dsqrt_rn(r2) < cutoff;
}

// Energy and force of one pair in range (ReferenceInteraxPairIxn.h:71-96): soft core, the pair terms, switching,
// alchemical scaling.  d = r_j - r_i; returns the force on i as fs*d.  ALCH: some particle of the handle is
// alchemical; without, lambda is 1 and there is no soft core, and the variant leaves those steps out.
template <typename Real, int TERMS, bool ALCH, typename Particle>
__device__ __forceinline__ void evaluate_pair(const TileArgs<Real>& a, const PairTables& tb, Real r2, const Particle& P, const Particle& Q,
                                              bool alchI, int flagJ, int origI, Real& e, Real& fs) {
    const Real invR = fast_rsqrt(r2);
    const Real r = r2*invR;
    Real rEff = r, invREff = invR;
    bool alchPair = false;
    Real chain = 1;   // (r/r_eff)^5, the derivative of r_eff, on soft-core pairs
    if constexpr (ALCH) {
        const bool alchJ = ((flagJ >> 30) & 1) != 0;
        alchPair = (alchI != alchJ);
        if (a.useSoftcore && alchPair) {
            const Real x = r2*r2*r2+a.softcoreTerm;
            if constexpr (std::is_same<Real, float>::value) {
                // x^(1/6): v_log_f32/v_exp_f32 estimate, then one Newton step on y^6 = x (error a few ulp)
                const float y = fast_exp2(__builtin_amdgcn_logf(x)*(1.0f/6.0f));
                const float y2 = y*y;
                rEff = y*(5.0f/6.0f) + x*fast_rcp(y2*y2*y)*(1.0f/6.0f);
                invREff = fast_rcp(rEff);
            }
            else {
                rEff = pow(x, Real(1.0/6.0));
                invREff = Real(1)/rEff;
            }
            if (rEff != r) {
                const Real ratio = r*invREff;
                const Real ratio2 = ratio*ratio;
                chain = ratio2*ratio2*ratio;
            }
        }
    }
    Real dEdR;
    const bool iFirst = origI < (flagJ & INDEX_MASK);
    if constexpr (std::is_same<Particle, FixedPair>::value)
        pair_energy_fixed<TERMS>(rEff, invREff, P, tb.beta2, e, dEdR);
    else if constexpr (std::is_same<Particle, TableParticle>::value)
        pair_energy_table<TERMS>(rEff, invREff, P, Q, iFirst, tb, e, dEdR);
    else
        pair_energy<Real, TERMS>(rEff, invREff, P, Q, iFirst, a.settings, e, dEdR);
    if constexpr (ALCH)
        dEdR *= chain;
    if (a.useSwitch && r > a.switchDistance) {
        Real x = (r-a.switchDistance)*a.invSwitchWidth;
        Real sw = 1+x*x*x*(Real(-10)+x*(Real(15)-x*Real(6)));
        Real dsw = x*x*(Real(-30)+x*(Real(60)-x*Real(30)))*a.invSwitchWidth;
        dEdR = dEdR*sw + e*dsw;
        e *= sw;
    }
    if constexpr (ALCH) {
        const Real scale = (alchPair ? a.lambda : Real(1));
        e *= scale;
        fs = scale*dEdR*invR;
    }
    else
        fs = dEdR*invR;
}

// The parameters of one slot, from global memory.
__device__ __forceinline__ void load_particle(const TileArgs<float>& a, int slot, TableParticle& P) {
    P.wd = a.pA[slot].x;
    P.cls = a.cls[slot];
}
template <typename Real>
__device__ __forceinline__ void load_particle(const TileArgs<Real>& a, int slot, PairParticle<Real>& P) {
    const typename Vec4T<Real>::type A = a.pA[slot], B = a.pB[slot], C = a.pC[slot];
    P.m1 = A.x; P.m2 = A.y; P.m3 = A.z; P.alpha = A.w;
    P.wd = B.x; P.amp = B.y; P.wx = B.z; P.iwx = B.w;
    P.k = C.x; P.l = C.y; P.ws = C.z; P.iws = C.w;
}

// Load the partner chunk of a tile into LDS, in the frame of block I.  Returns the partner's flags (-1 for none).
template <typename Real, bool TABLES>
__device__ __forceinline__ int stage_partners(const TileArgs<Real>& a, int2 tile, int lane, SideLds<Real, TABLES>& s) {
    typedef typename Vec4T<Real>::type R4;
    const int I = tile.x & TILE_BLOCK_MASK;
    const int slot = (tile.y < 0 ? I*BLOCK+lane : a.partners[tile.y+lane]);
    if (slot < 0) {
        R4 zero = {0, 0, 0, 0};
        store_pos(s, lane, zero);
        s.flags[lane] = -1;
        return -1;
    }
    const int J = slot/BLOCK;
    R4 pj = a.local[slot];
    if (J != I) {
        double sx = a.centers[3*J]-a.centers[3*I], sy = a.centers[3*J+1]-a.centers[3*I+1], sz = a.centers[3*J+2]-a.centers[3*I+2];
        if (a.periodic)
            reference_min_image(sx, sy, sz, a.box);
        pj.x += Real(sx);
        pj.y += Real(sy);
        pj.z += Real(sz);
    }
    store_pos(s, lane, pj);
    return stage_particle<Real, TABLES>(s, lane, a, slot, decode_index(pj.w));
}

// Energy and pair count of a wave, written once at its end: the launch's totals are then summed in wave order.
// pairs: this lane's count, or with waveTotal the wave's.
__device__ __forceinline__ void store_wave_totals(double* waveEnergy, int* waveCount, long long wave, int lane, double energy, int pairs,
                                                  bool waveTotal = false) {
    const double e = wave_sum(energy);
    const int c = (waveTotal ? pairs : (int) wave_sum((double) pairs));
    if (lane == 0) {
        waveEnergy[wave] = e;
        waveCount[wave] = c;
    }
}

// Tiles without a cutoff: every candidate interacts.  64 rotation steps, lane l paired with j = (l+s) % 64.
template <typename Real, int TERMS, bool TABLES, bool ALCH>
__global__ void __launch_bounds__(256) k_pair_rotate(TileArgs<Real> a) {
    typedef typename Vec4T<Real>::type R4;
    typedef typename ParticleType<Real, TABLES>::type Particle;
    __shared__ SideLds<Real, TABLES> side[4];
    extern __shared__ float tableLds[];   // the type tables, sized at launch
    PairTables tb = {};
    if constexpr (TABLES)
        tb = load_tables<Real>(a, tableLds);
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    SideLds<Real, TABLES>& s = side[w];
    const long long wave = (long long) blockIdx.x*4+w;
    const long long numWaves = (long long) gridDim.x*4;
    const long long T = a.numTiles;
    const long long begin = T*wave/numWaves, end = T*(wave+1)/numWaves;

    int curI = -1;
    R4 pi = {0, 0, 0, 0};
    int flagI = -1;
    long long fixI[3] = {0, 0, 0};
    Particle P, Q;
    double energy = 0;
    int count = 0;

    for (long long item = begin; item < end; item++) {
        const long long t = (a.tileIndex != nullptr ? a.tileIndex[item] : item);
        const int2 tile = a.tiles[t];
        const int I = tile.x & TILE_BLOCK_MASK;
        if (I != curI) {
            if (curI >= 0 && flagI >= 0) {
                int oi = flagI & INDEX_MASK;
                add_fixed(&a.force[oi], fixI[0]);
                add_fixed(&a.force[a.stride+oi], fixI[1]);
                add_fixed(&a.force[2*a.stride+oi], fixI[2]);
            }
            fixI[0] = fixI[1] = fixI[2] = 0;
            curI = I;
            int slot = I*BLOCK+lane;
            pi = a.local[slot];
            flagI = a.flags[slot];
            load_particle(a, slot, P);
        }
        const int flagsJ = stage_partners<Real, TABLES>(a, tile, lane, s);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const unsigned long long mrow = a.masks[(size_t) t*BLOCK+lane];
        const bool diag = (tile.y < 0);
        const int sBegin = (diag ? 1 : 0), sEnd = (diag ? 33 : 64);
        const bool alchI = (flagI >= 0 && ((flagI >> 30) & 1) != 0);
        const int origI = decode_index(pi.w);
        Real fi[3] = {0, 0, 0};
        Real fj[3] = {0, 0, 0};
        double tileE = 0;
        for (int st = sBegin; st < sEnd; st++) {
            const int jl = (lane+st) & 63;
            bool interact = ((mrow >> jl) & 1ULL) != 0;
            if (diag && st == 32 && lane >= 32)
                interact = false;
            if (interact) {
                const R4 pj = load_pos(s, jl);
                const Real dx = pj.x-pi.x, dy = pj.y-pi.y, dz = pj.z-pi.z;
                const Real r2 = dx*dx+dy*dy+dz*dz;
                const int flagJ = s.flags[jl];
                fetch_particle(s, jl, Q);
                Real e, fs;
                evaluate_pair<Real, TERMS, ALCH>(a, tb, r2, P, Q, alchI, flagJ, origI, e, fs);
                tileE += (double) e;
                fi[0] += dx*fs;
                fi[1] += dy*fs;
                fi[2] += dz*fs;
                fj[0] -= dx*fs;
                fj[1] -= dy*fs;
                fj[2] -= dz*fs;
                count++;
            }
            fj[0] = wave_rotate_down(fj[0]);
            fj[1] = wave_rotate_down(fj[1]);
            fj[2] = wave_rotate_down(fj[2]);
        }
        // lane l now holds the accumulator of j = (l + sEnd) % 64, whose flags lane jl holds
        {
            const int jl = (lane+sEnd) & 63;
            const int flagJ = __shfl(flagsJ, jl, 64);
            if (flagJ >= 0) {
                int oj = flagJ & INDEX_MASK;
                add_fixed(&a.force[oj], to_fixed(fj[0]));
                add_fixed(&a.force[a.stride+oj], to_fixed(fj[1]));
                add_fixed(&a.force[2*a.stride+oj], to_fixed(fj[2]));
            }
        }
        fixI[0] += to_fixed(fi[0]);
        fixI[1] += to_fixed(fi[1]);
        fixI[2] += to_fixed(fi[2]);
        energy += tileE;
        if (a.tileEnergy != nullptr) {
            const double te = wave_sum(tileE);
            if (lane == 0)
                a.tileEnergy[t] = te;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (curI >= 0 && flagI >= 0) {
        int oi = flagI & INDEX_MASK;
        add_fixed(&a.force[oi], fixI[0]);
        add_fixed(&a.force[a.stride+oi], fixI[1]);
        add_fixed(&a.force[2*a.stride+oi], fixI[2]);
    }
    store_wave_totals(a.waveEnergy, a.waveCount, wave, lane, energy, count);
}

// Tiles with a cutoff.  A lockstep filter pass over the 64 rotation steps pushes the pairs inside the cutoff into a
// per-wave ring queue in LDS (ballot + rank, so the order is deterministic); every full batch of 64 queued pairs is
// then evaluated with all lanes busy, whatever the spread of neighbour counts between the particles of the tile.
// The forces on i and j are added in LDS as 64-bit fixed point (the j side gets the exact negation of the i side's
// value) and flushed to the force buffer per tile (j) and per work item (i).  (LDS FP32 atomics, ds_add_f32, take
// about 190 cycles per wave instruction on gfx950 against 7 for ds_add_u64, so the forces are converted per pair.)
constexpr int QUEUE = 512;   // pending pairs never exceed 63 + 256

template <typename Real, bool TABLES>
struct QueueLds {
    SideLds<Real, TABLES> i, j;
    unsigned long long fI[3][BLOCK], fJ[3][BLOCK];
    unsigned short queue[QUEUE];
};

template <typename Real, int TERMS, bool TABLES, bool ALCH, bool FIXED>
__device__ __forceinline__ void evaluate_batch(const TileArgs<Real>& a, const PairTables& tb, QueueLds<Real, TABLES>& q, int head,
                                               int valid, int lane, bool pairImage, double& energy) {
    typedef typename Vec4T<Real>::type R4;
    typedef typename ParticleType<Real, TABLES>::type Particle;
    if (lane >= valid)
        return;
    const unsigned e = q.queue[(head+lane) & (QUEUE-1)];
    const int il = (int) (e >> 6), jl = (int) (e & 63);
    const R4 pi = load_pos(q.i, il), pj = load_pos(q.j, jl);
    Real dx = pj.x-pi.x, dy = pj.y-pi.y, dz = pj.z-pi.z;
    if (pairImage)
        min_image<Real>(dx, dy, dz, a.boxR[0], a.boxR[1], a.boxR[2], a.boxR[3], a.boxR[4], a.boxR[5], a.boxR[6], a.boxR[7],
                        a.boxR[8]);
    const Real r2 = dx*dx+dy*dy+dz*dz;
    // the flags hold the alchemical bits and the System indices, which only an InteraxOrder2 overlap needs (for the
    // Exact scheme the overlap of a pair is the same either way round)
    int flagI = 0, flagJ = 0;
    if (ALCH || (!FIXED && a.ordered)) {
        flagI = q.i.flags[il];
        flagJ = q.j.flags[jl];
    }
    Real en, fs;
    if constexpr (FIXED) {
        // the pair's rows from the sum of the two role words (fixed_role_word)
        const int word = __float_as_int(q.i.pw[il])+__float_as_int(q.j.pw[jl]);
        const char* base = (const char*) tb.disp;
        FixedPair P;
        P.disp = *(const float4*) (base+(word & 0xffff));
        P.row = (const OverlapRow*) (base+(word >> 16));
        P.wsum = q.i.wd[il]+q.j.wd[jl];
        evaluate_pair<Real, TERMS, ALCH>(a, tb, r2, P, P, ((flagI >> 30) & 1) != 0, flagJ, flagI & INDEX_MASK, en, fs);
    }
    else {
        Particle P, Q;
        fetch_particle(q.i, il, P);
        fetch_particle(q.j, jl, Q);
        evaluate_pair<Real, TERMS, ALCH>(a, tb, r2, P, Q, ((flagI >> 30) & 1) != 0, flagJ, flagI & INDEX_MASK, en, fs);
    }
    energy += (double) en;
    const long long fx = to_fixed(dx*fs), fy = to_fixed(dy*fs), fz = to_fixed(dz*fs);
    atomicAdd(&q.fI[0][il], (unsigned long long) fx);
    atomicAdd(&q.fI[1][il], (unsigned long long) fy);
    atomicAdd(&q.fI[2][il], (unsigned long long) fz);
    // the j accumulators sum the forces on i: their exact negation is added at the flush
    atomicAdd(&q.fJ[0][jl], (unsigned long long) fx);
    atomicAdd(&q.fJ[1][jl], (unsigned long long) fy);
    atomicAdd(&q.fJ[2][jl], (unsigned long long) fz);
}

// One lane's partner particle of a tile, loaded from global memory.
template <typename Real, bool TABLES>
struct PartnerFetch {
    typedef typename Vec4T<Real>::type R4;
    R4 pos, A, B, C;
    int cls, flags, block;
    double cx, cy, cz;   // origin of the partner's block frame
};

template <typename Real, bool TABLES>
__device__ __forceinline__ void fetch_partner(const TileArgs<Real>& a, int slot, PartnerFetch<Real, TABLES>& f) {
    if (slot < 0) {
        f.flags = -1;
        f.block = -1;
        return;
    }
    f.block = slot/BLOCK;
    f.pos = a.local[slot];
    f.A = a.pA[slot];
    if constexpr (TABLES)
        f.cls = a.cls[slot];
    else {
        f.B = a.pB[slot];
        f.C = a.pC[slot];
    }
    f.flags = a.flags[slot];
    f.cx = a.centers[3*f.block];
    f.cy = a.centers[3*f.block+1];
    f.cz = a.centers[3*f.block+2];
}

// Write a prefetched partner into LDS, moved into the frame of block I (origin cI).
template <typename Real, bool TABLES, int TERMS = 0, bool FIXED = false>
__device__ __forceinline__ void store_partner(const TileArgs<Real>& a, const PartnerFetch<Real, TABLES>& f, int I, double cIx, double cIy,
                                              double cIz, int lane, SideLds<Real, TABLES>& s) {
    typedef typename Vec4T<Real>::type R4;
    if (f.flags < 0) {
        R4 zero = {0, 0, 0, 0};
        store_pos(s, lane, zero);
        s.flags[lane] = -1;
        return;
    }
    R4 pj = f.pos;
    if (f.block != I) {
        double sx = f.cx-cIx, sy = f.cy-cIy, sz = f.cz-cIz;
        if (a.periodic)
            min_image<double>(sx, sy, sz, a.box.ax, a.box.bx, a.box.by, a.box.cx, a.box.cy, a.box.cz, a.boxInv[0], a.boxInv[1],
                              a.boxInv[2]);
        pj.x += Real(sx);
        pj.y += Real(sy);
        pj.z += Real(sz);
    }
    if constexpr (TABLES) {
        if constexpr (FIXED)
            pj.w = __int_as_float(fixed_role_word<TERMS>(a, f.cls, false));
        else
            pj.w = __int_as_float(f.cls);
        s.wd[lane] = f.A.x;
    }
    else {
        s.A[lane] = f.A;
        s.B[lane] = f.B;
        s.C[lane] = f.C;
    }
    store_pos(s, lane, pj);
    s.flags[lane] = lds_flags(f.flags, decode_index(f.pos.w));
}

// With type tables the kernel is held to 5 waves per SIMD (at most 102 VGPRs; a few values spill): the pair batches
// are latency-bound, and the extra wave hides more of it than the spills cost.  A workgroup is QUEUE_GROUP_WAVES
// wavefronts (measured on the 98 304-atom box: 1 wave, 185.5 us, but four times the partial sums to reduce, 0.209 ms
// per step; 2 waves, 207 us; 4 waves, 186-190 us and 0.205 ms per step).
constexpr int QUEUE_GROUP_WAVES = 4;
// FIXED: the fixed table form (pair_energy_fixed), for tables that allow it.
template <typename Real, int TERMS, bool TABLES, bool ALCH, bool FIXED = false>
__global__ void __launch_bounds__(64*QUEUE_GROUP_WAVES, TABLES ? 5 : 1) k_pair_queue(TileArgs<Real> a) {
    typedef typename Vec4T<Real>::type R4;
    __shared__ QueueLds<Real, TABLES> lds[QUEUE_GROUP_WAVES];
    extern __shared__ float tableLds[];   // the type tables, sized at launch
    PairTables tb = {};
    if constexpr (TABLES)
        tb = load_tables<Real>(a, tableLds);
    const int lane = threadIdx.x & 63;
    // the wavefront's index as a wave-uniform value (the work item and tile indices derived from it are then uniform)
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    QueueLds<Real, TABLES>& q = lds[w];
    const long long wave = (long long) blockIdx.x*QUEUE_GROUP_WAVES+w;
    const long long numWaves = (long long) gridDim.x*QUEUE_GROUP_WAVES;
    const long long T = a.numTiles;
    long long begin = T*wave/numWaves, end = T*(wave+1)/numWaves;
    if (a.items != nullptr) {
        const int2 item = a.items[wave];
        begin = item.x;
        end = item.y;
    }
    const Real margin = a.band*a.cutoff2;
    const Real inner = a.cutoff2-margin, outer = a.cutoff2+margin;

    int curI = -1;
    R4 pi = {0, 0, 0, 0};
    int flagI = -1;
    double cIx = 0, cIy = 0, cIz = 0;
    double energy = 0;
    int pairs = 0;

    // A tile's partners and first candidate entries are loaded before the previous tile's j-side flush: vector memory
    // operations are counted in order, so a load issued after the flush's atomics could only be waited for together
    // with them; issued first, they are in flight together with the flush.
    long long t = 0;
    int tileX = 0, cBegin = 0, cEnd = 0;
    PartnerFetch<Real, TABLES> f;
    uint2 firstChunk = {0, 0};
    bool firstValid = false;
    auto prefetch_tile = [&](long long item) {
        t = (a.tileIndex != nullptr ? a.tileIndex[item] : item);
        // the tile word, the lane's partner slot and the list bounds together, then what depends on them
        tileX = a.tiles[t].x;
        const int slot = a.tileSlots[(size_t) t*BLOCK+lane];
        cBegin = a.listOffsets[t];
        cEnd = a.listOffsets[t+1];
        // every lane loads (from the tile's last four entries, or the start of the array, when past the end) and keeps
        // a validity flag; offsets are multiples of 4 and the array holds at least 4
        firstValid = (cBegin+4*lane < cEnd);
        firstChunk = *(const uint2*) (a.lists+(firstValid ? cBegin+4*lane : (cEnd >= 4 ? cEnd-4 : 0)));
        fetch_partner<Real, TABLES>(a, slot, f);
    };
    if (begin < end)
        prefetch_tile(begin);
    for (long long item = begin; item < end; item++) {
        const int I = tileX & TILE_BLOCK_MASK;
        const bool pairImage = (tileX & TILE_PAIR_IMAGE) != 0;
        if (I != curI) {
            if (curI >= 0 && flagI >= 0) {
                int oi = flagI & INDEX_MASK;
                add_fixed(&a.force[oi], (long long) q.fI[0][lane]);
                add_fixed(&a.force[a.stride+oi], (long long) q.fI[1][lane]);
                add_fixed(&a.force[2*a.stride+oi], (long long) q.fI[2][lane]);
            }
            curI = I;
            const int slot = I*BLOCK+lane;
            pi = a.local[slot];
            flagI = a.flags[slot];
            cIx = a.centers[3*I];
            cIy = a.centers[3*I+1];
            cIz = a.centers[3*I+2];
            store_pos(q.i, lane, pi);
            stage_particle<Real, TABLES, TERMS, FIXED>(q.i, lane, a, slot, decode_index(pi.w));
            q.fI[0][lane] = q.fI[1][lane] = q.fI[2][lane] = 0;
        }
        store_partner<Real, TABLES, TERMS, FIXED>(a, f, I, cIx, cIy, cIz, lane, q.j);
        q.fJ[0][lane] = q.fJ[1][lane] = q.fJ[2][lane] = 0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        // The tile's candidate list, 256 entries (4 per lane) per step with the next step's entries in flight (each
        // chunk stored transposed: sub-step s of lane l is entry s m + l of the chunk's rotation order): each
        // entry in range is pushed into the queue, then every full batch of 64 queued pairs is evaluated.  The four
        // filter sub-steps are unrolled and hold no batch evaluation, so no loop inside a step makes the compiler
        // wait for the next step's entries before the filter starts.
        // Every lane loads (from the tile's last four entries, or the start of the array, when past the end) and
        // keeps a validity flag instead of branching, so that the wait for a step's entries can count the next
        // step's load as outstanding on every path.  Offsets are multiples of 4 and the array holds at least 4.
        const int cLast = (cEnd >= 4 ? cEnd-4 : 0);
        int head = 0, count = 0;
        bool chunkValid = firstValid;
        uint2 chunk = firstChunk;
        for (int base = cBegin; base < cEnd; base += 4*BLOCK) {
            const int nextIndex = base+4*BLOCK+4*lane;
            const bool nextValid = (nextIndex < cEnd);
            const uint2 next = *(const uint2*) (a.lists+(nextValid ? nextIndex : cLast));
#pragma unroll
            for (int sub = 0; sub < 4; sub++) {
                const unsigned word = (sub < 2 ? chunk.x : chunk.y);
                const unsigned e = (word >> ((sub & 1)*16)) & 0xffffu;
                bool in = chunkValid && (e != 0xffffu);
                if (in) {
                    const int il = (int) (e >> 6), jl = (int) (e & 63);
                    const R4 pa = load_pos(q.i, il), pb = load_pos(q.j, jl);
                    Real dx = pb.x-pa.x, dy = pb.y-pa.y, dz = pb.z-pa.z;
                    if (pairImage)
                        min_image<Real>(dx, dy, dz, a.boxR[0], a.boxR[1], a.boxR[2], a.boxR[3], a.boxR[4], a.boxR[5], a.boxR[6],
                                        a.boxR[7], a.boxR[8]);
                    const Real r2 = dx*dx+dy*dy+dz*dz;
                    if (r2 >= inner)
                        in = (r2 < outer) && exact_within_cutoff(a.posq, a.correction, a.precision, a.periodic, a.box, a.cutoff,
                                                                 a.perm[I*BLOCK+il], a.perm[a.tileSlots[(size_t) t*BLOCK+jl]]);
                }
                const unsigned long long ballot = wave_ballot(in);
                if (in)
                    q.queue[lanes_below(ballot, (unsigned) (head+count)) & (QUEUE-1)] = (unsigned short) e;
                count += __popcll(ballot);
            }
            while (count >= BLOCK) {
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                evaluate_batch<Real, TERMS, TABLES, ALCH, FIXED>(a, tb, q, head, BLOCK, lane, pairImage, energy);
                pairs += BLOCK;
                head = (head+BLOCK) & (QUEUE-1);
                count -= BLOCK;
            }
            chunk = next;
            chunkValid = nextValid;
        }
        if (count > 0) {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            evaluate_batch<Real, TERMS, TABLES, ALCH, FIXED>(a, tb, q, head, count, lane, pairImage, energy);
            pairs += count;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        {
            const int flagJ = f.flags;
            if (item+1 < end)
                prefetch_tile(item+1);
            if (flagJ >= 0) {
                int oj = flagJ & INDEX_MASK;
                add_fixed(&a.force[oj], -(long long) q.fJ[0][lane]);
                add_fixed(&a.force[a.stride+oj], -(long long) q.fJ[1][lane]);
                add_fixed(&a.force[2*a.stride+oj], -(long long) q.fJ[2][lane]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (curI >= 0 && flagI >= 0) {
        int oi = flagI & INDEX_MASK;
        add_fixed(&a.force[oi], (long long) q.fI[0][lane]);
        add_fixed(&a.force[a.stride+oi], (long long) q.fI[1][lane]);
        add_fixed(&a.force[2*a.stride+oi], (long long) q.fI[2][lane]);
    }
    // one energy and pair count per workgroup, its waves summed in wave order
    __shared__ double groupEnergy[QUEUE_GROUP_WAVES];
    __shared__ int groupPairs[QUEUE_GROUP_WAVES];
    const double waveTotal = wave_sum(energy);
    if (lane == 0) {
        groupEnergy[w] = waveTotal;
        groupPairs[w] = pairs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double e = 0;
        int c = 0;
        for (int k = 0; k < QUEUE_GROUP_WAVES; k++) {
            e += groupEnergy[k];
            c += groupPairs[k];
        }
        a.waveEnergy[blockIdx.x] = e;
        a.waveCount[blockIdx.x] = c;
    }
}

// Sum the per-wave energies and pair counts of a launch (one workgroup, a fixed assignment of partials to threads
// and a fixed reduction tree, so the result does not depend on timing) and add the energy to the energy buffer.
__global__ void __launch_bounds__(1024) k_reduce_waves(const double* waveEnergy, const int* waveCount, int numWaves, double* energy,
                                                       long long* pairCount, int addEnergy) {
    __shared__ double sE[1024];
    __shared__ long long sC[1024];
    // strided (coalesced) partial sums, four loads in flight per thread
    double e = 0;
    long long c = 0;
    for (int k0 = threadIdx.x; k0 < numWaves; k0 += 4*1024) {
        double v[4] = {0, 0, 0, 0};
        int n[4] = {0, 0, 0, 0};
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (k0+u*1024 < numWaves) {
                v[u] = waveEnergy[k0+u*1024];
                n[u] = waveCount[k0+u*1024];
            }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            e += v[u];
            c += n[u];
        }
    }
    sE[threadIdx.x] = e;
    sC[threadIdx.x] = c;
    __syncthreads();
    for (int stride = 512; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) {
            sE[threadIdx.x] += sE[threadIdx.x+stride];
            sC[threadIdx.x] += sC[threadIdx.x+stride];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (addEnergy)
            energy[0] += sE[0];
        *pairCount = sC[0];
    }
}

// Energy of every segment: the sum of its tiles' energies in tile order (one workgroup per segment, fixed tree),
// added to energies[s].
__global__ void __launch_bounds__(256) k_reduce_segments(const double* tileEnergy, const int* segTiles, const int* segList,
                                                          double* energies) {
    __shared__ double sE[256];
    const int seg = (segList != nullptr ? segList[blockIdx.x] : blockIdx.x);
    const int begin = segTiles[seg], end = segTiles[seg+1];
    double e = 0;
    for (int t = begin+threadIdx.x; t < end; t += 256)
        e += tileEnergy[t];
    sE[threadIdx.x] = e;
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride)
            sE[threadIdx.x] += sE[threadIdx.x+stride];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        energies[seg] += sE[0];
}

// ------------------------------------------------------------------------------------------------------------
// Long range correction (InteraxDispersionForceImpl.cpp:86-149), FP64 on the device
// ------------------------------------------------------------------------------------------------------------

struct LrcArgs {
    const double* classes;   // 7 per class: m1, m2, m3, alpha, width, alchemical, count
    const long long* numClasses;   // device
    double cutoff, switchDistance, switchWidth, dampingScale;
    int useSwitch;
    double nodes[32], weights[32];
};

__device__ double lrc_pair_energy(double r, const double* c1, const double* c2, double beta) {
    PairParticle<double> P, Q;
    P.m1 = c1[0]; P.m2 = c1[1]; P.m3 = c1[2]; P.alpha = c1[3]; P.wd = c1[4];
    Q.m1 = c2[0]; Q.m2 = c2[1]; Q.m3 = c2[2]; Q.alpha = c2[3]; Q.wd = c2[4];
    PairSettings set;
    set.damping_scale = beta;
    set.damping_scale_f = (float) beta;
    set.shared_overlap = 0;
    double e, d;
    pair_energy<double, IXH_TERM_DISPERSION>(r, 1.0/r, P, Q, true, set, e, d);
    return e;
}

constexpr int LRC_THREADS = 256, LRC_BLOCKS = 2048;

__global__ void __launch_bounds__(LRC_THREADS) k_long_range(LrcArgs a, double* partial /* 2 per workgroup */) {
    long long tid = (long long) blockIdx.x*blockDim.x+threadIdx.x;
    long long nthreads = (long long) gridDim.x*blockDim.x;
    long long C = *a.numClasses;
    long long totalPairs = C*(C+1)/2;
    long long begin = totalPairs*tid/nthreads, end = totalPairs*(tid+1)/nthreads;
    double sum = 0, alch = 0;
    if (begin < end) {
        // decode begin into (i, j), pairs ordered i <= j, row i has C - i entries
        long long i = 0, rowStart = 0;
        {
            // solve rowStart(i) = i*C - i(i-1)/2 <= begin
            double cc = (double) C;
            double guess = floor(((2*cc+1)-sqrt((2*cc+1)*(2*cc+1)-8.0*(double) begin))/2);
            i = (long long) fmax(0.0, guess-2);
            while (i+1 < C && (i+1)*C-(i+1)*i/2 <= begin)
                i++;
            rowStart = i*C-i*(i-1)/2;
        }
        long long j = i+(begin-rowStart);
        for (long long p = begin; p < end; p++) {
            const double* c1 = a.classes+7*i;
            const double* c2 = a.classes+7*j;
            if (c1[3]*c2[3] != 0) {
                double integral = 0;
                for (int k = 0; k < 32; k++) {
                    double u = a.nodes[k];
                    integral += a.weights[k]*lrc_pair_energy(a.cutoff/u, c1, c2, a.dampingScale)*a.cutoff*a.cutoff*a.cutoff/(u*u*u*u);
                    if (a.useSwitch) {
                        double x = a.nodes[k];
                        double r = a.switchDistance+x*a.switchWidth;
                        double oneMinusSwitch = x*x*x*(10+x*(-15+x*6));
                        integral += a.weights[k]*a.switchWidth*oneMinusSwitch*lrc_pair_energy(r, c1, c2, a.dampingScale)*r*r;
                    }
                }
                double count = (i == j ? 0.5*c1[6]*(c1[6]+1) : c1[6]*c2[6]);
                if (c1[5] != 0 || c2[5] != 0)
                    alch += count*integral;
                else
                    sum += count*integral;
            }
            j++;
            if (j >= C) {
                i++;
                j = i;
            }
        }
    }
    // workgroup sums in a fixed order
    __shared__ double sSum[LRC_THREADS/64], sAlch[LRC_THREADS/64];
    sum = wave_sum(sum);
    alch = wave_sum(alch);
    if (threadIdx.x%64 == 0) {
        sSum[threadIdx.x/64] = sum;
        sAlch[threadIdx.x/64] = alch;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0, t = 0;
        for (int w = 0; w < LRC_THREADS/64; w++) {
            s += sSum[w];
            t += sAlch[w];
        }
        partial[2*blockIdx.x] = s;
        partial[2*blockIdx.x+1] = t;
    }
}

void gauss_legendre(int n, double* nodes, double* weights) {
    // InteraxDispersionForceImpl.cpp:49-73
    for (int i = 0; i < n; i++) {
        double x = std::cos(M_PI*(i+0.75)/(n+0.5));
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
            if (std::fabs(step) < 1e-15)
                break;
        }
        nodes[i] = 0.5*(1-x);
        weights[i] = 1/((1-x*x)*derivative*derivative);
    }
}

BoxD make_box(const double* box) {
    BoxD b;
    b.ax = box[0];
    b.bx = box[3];
    b.by = box[4];
    b.cx = box[6];
    b.cy = box[7];
    b.cz = box[8];
    return b;
}

int grid_for(long long n, int threads) {
    return (int) std::max(1LL, (n+threads-1)/threads);
}

// ------------------------------------------------------------------------------------------------------------
// Parameter pipeline: every parameter update, from host arrays (ixh_pair_set_parameters) or device arrays
// (ixh_pair_set_parameters_device), lands in FP64 device arrays in the reference's per-particle order, from which
// these kernels derive the evaluation's arrays, check the type-table classes, and compute the long range correction
// with the classes grouped on the device.
// ------------------------------------------------------------------------------------------------------------

struct DeriveArgs {
    const double* disp;      // 5 per particle (m1, m2, m3, alpha, width), or nullptr
    const double* exch;      // 2 per particle (amplitude, width), or nullptr
    const double* sr;        // 3 per particle (k, l, width), or nullptr
    int n;
    int precision;
    int useTables;
    int shared;              // one overlap for exchange and short range (both present): their widths must be equal
    const int* cls;          // type tables: dispersion | exchange << 8 | short range << 16 class of every particle
    const int* reps;         // type tables: the first particle of every dispersion, exchange, short-range class
    int repExch, repSr;      // offsets of the exchange and short-range classes in reps
    void* A;                 // the evaluation's per-particle vectors (float4 or double4), particle order
    void* B;
    void* C;
    double* lambdaParams;    // 12 per particle (PairParticle order)
    int* mismatch;           // type tables: set when a particle no longer has the parameters of its class's first one
};

__global__ void k_derive_params(DeriveArgs a) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= a.n)
        return;
    double A[4] = {0, 0, 0, 0}, B[4] = {0, 0, 0, 0}, C[4] = {0, 0, 0, 0};
    if (a.disp != nullptr) {
        for (int k = 0; k < 4; k++)
            A[k] = a.disp[5*(size_t) i+k];
        B[0] = a.disp[5*(size_t) i+4];
    }
    if (a.exch != nullptr) {
        B[1] = a.exch[2*(size_t) i];
        B[2] = a.exch[2*(size_t) i+1];
        B[3] = (B[2] != 0 ? 1.0/B[2] : 0.0);
    }
    if (a.sr != nullptr) {
        for (int k = 0; k < 3; k++)
            C[k] = a.sr[3*(size_t) i+k];
        C[3] = (C[2] != 0 ? 1.0/C[2] : 0.0);
    }
    double* q = a.lambdaParams+12*(size_t) i;
    for (int k = 0; k < 4; k++) {
        q[k] = A[k];
        q[4+k] = B[k];
        q[8+k] = C[k];
    }
    // one overlap for both terms needs equal widths; with type tables, every particle must still equal its class's
    // first particle in every parameter the class stands for
    bool same = (!a.shared || B[2] == C[2]);
    if (a.useTables) {
        const int c = a.cls[i];
        if (a.disp != nullptr) {
            const int r = a.reps[c & 0xff];
            for (int k = 0; k < 4; k++)
                same = same && (a.disp[5*(size_t) r+k] == A[k]);
        }
        if (a.exch != nullptr) {
            const int r = a.reps[a.repExch+((c >> 8) & 0xff)];
            same = same && a.exch[2*(size_t) r] == B[1] && a.exch[2*(size_t) r+1] == B[2];
            if (a.shared)
                same = same && a.sr[3*(size_t) r] == C[0] && a.sr[3*(size_t) r+1] == C[1];
        }
        if (a.sr != nullptr && !a.shared) {
            const int r = a.reps[a.repSr+((c >> 16) & 0xff)];
            for (int k = 0; k < 3; k++)
                same = same && (a.sr[3*(size_t) r+k] == C[k]);
        }
    }
    if (!same)
        *a.mismatch = 1;
    if (a.precision == IXH_DOUBLE) {
        ((double4*) a.A)[i] = make_double4(A[0], A[1], A[2], A[3]);
        ((double4*) a.B)[i] = make_double4(B[0], B[1], B[2], B[3]);
        ((double4*) a.C)[i] = make_double4(C[0], C[1], C[2], C[3]);
        return;
    }
    // with type tables the first vector holds the dispersion width only (the classes carry the rest)
    if (a.useTables)
        ((float4*) a.A)[i] = make_float4((float) B[0], 0.0f, 0.0f, 0.0f);
    else
        ((float4*) a.A)[i] = make_float4((float) A[0], (float) A[1], (float) A[2], (float) A[3]);
    ((float4*) a.B)[i] = make_float4((float) B[0], (float) B[1], (float) B[2], (float) B[3]);
    ((float4*) a.C)[i] = make_float4((float) C[0], (float) C[1], (float) C[2], (float) C[3]);
}

__global__ void k_clear_int(int* p) {
    *p = 0;
}

// The FP64 parameters of the class representatives (10 per representative: 5 dispersion, 2 exchange, 3 short range),
// from which the host forms the type tables.
__global__ void k_gather_reps(const double* disp, const double* exch, const double* sr, const int* reps, int count, double* out) {
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= count)
        return;
    const int i = reps[k];
    double* o = out+10*k;
    for (int c = 0; c < 5; c++)
        o[c] = (disp != nullptr ? disp[5*(size_t) i+c] : 0.0);
    for (int c = 0; c < 2; c++)
        o[5+c] = (exch != nullptr ? exch[2*(size_t) i+c] : 0.0);
    for (int c = 0; c < 3; c++)
        o[7+c] = (sr != nullptr ? sr[3*(size_t) i+c] : 0.0);
}

// Dispersion classes of the long range correction: particles of identical (m1, m2, m3, alpha, width, alchemical), in
// the lexicographic order of that tuple (InteraxDispersionForceImpl.cpp:94-101), found on the device by six stable
// radix sorts of the particle order, least significant field first.  A double's bits, with the sign bit flipped (or
// all bits, for a negative value), sort as the value does; -0.0 is taken as +0.0.
__device__ __forceinline__ double class_field(const double* disp, const int* flags, int i, int field) {
    return (field < 5 ? disp[5*(size_t) i+field] : ((flags[i] & ALCHEMICAL_BIT) != 0 ? 1.0 : 0.0));
}

__global__ void k_lrc_keys(const double* disp, const int* flags, int n, int field, const int* perm, unsigned long long* keys) {
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= n)
        return;
    double x = class_field(disp, flags, perm[k], field);
    if (x == 0.0)
        x = 0.0;
    const unsigned long long bits = (unsigned long long) __double_as_longlong(x);
    keys[k] = ((bits >> 63) != 0 ? ~bits : bits | 0x8000000000000000ULL);
}

__global__ void k_iota(int* p, int n) {
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k < n)
        p[k] = k;
}

// first[k] = 1 where sorted position k starts a class
__global__ void k_lrc_bounds(const double* disp, const int* flags, const int* perm, int n, int* first) {
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= n)
        return;
    bool start = (k == 0);
    for (int f = 0; f < 6 && !start; f++)
        start = class_field(disp, flags, perm[k], f) != class_field(disp, flags, perm[k-1], f);
    first[k] = start ? 1 : 0;
}

// index[k]: the inclusive prefix sum of first[]: class index[k] - 1 holds sorted position k.  Writes the first sorted
// position of every class, the class count and classStart[count] = n.
__global__ void k_lrc_starts(const int* first, const int* index, int n, int* classStart, long long* numClasses) {
    const int k = blockIdx.x*blockDim.x+threadIdx.x;
    if (k >= n)
        return;
    if (first[k])
        classStart[index[k]-1] = k;
    if (k == n-1) {
        *numClasses = index[k];
        classStart[index[k]] = n;
    }
}

// 7 values per class: m1, m2, m3, alpha, width, alchemical, count
__global__ void k_lrc_classes(const double* disp, const int* flags, const int* perm, const int* classStart,
                              const long long* numClasses, double* classes) {
    const int c = blockIdx.x*blockDim.x+threadIdx.x;
    if (c >= *numClasses)
        return;
    const int i = perm[classStart[c]];
    for (int f = 0; f < 6; f++)
        classes[7*(size_t) c+f] = class_field(disp, flags, i, f);
    classes[7*(size_t) c+6] = (double) (classStart[c+1]-classStart[c]);
}

// The correction's coefficients from the partial sums of k_long_range (two per thread), summed in a fixed order:
// out = normalization (sum, alchemical sum) (InteraxDispersionForceImpl.cpp:145-148).
__global__ void __launch_bounds__(1024) k_lrc_reduce(const double* partial, int threads, double normalization, double* out) {
    __shared__ double sS[1024], sA[1024];
    double sum = 0, alch = 0;
    for (int t = threadIdx.x; t < threads; t += 1024) {
        sum += partial[2*(size_t) t];
        alch += partial[2*(size_t) t+1];
    }
    sS[threadIdx.x] = sum;
    sA[threadIdx.x] = alch;
    __syncthreads();
    for (int stride = 512; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) {
            sS[threadIdx.x] += sS[threadIdx.x+stride];
            sA[threadIdx.x] += sA[threadIdx.x+stride];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = normalization*sS[0];
        out[1] = normalization*sA[0];
    }
}

// ------------------------------------------------------------------------------------------------------------
// Energies at several lambda values (no counterpart in the reference)
// ------------------------------------------------------------------------------------------------------------

constexpr int MAX_LAMBDA_WINDOWS = 16;

struct LambdaArgs {
    const void* posq;
    const float4* correction;
    int precision;
    int n;
    const double* params;       // 12 per particle, in PairParticle order, original particle order
    const int* flags;           // particle index | alchemical bit
    const int* atomIndex;       // System index of each particle (the platform's atom index array), or nullptr
    const int* alchList;        // the alchemical particles
    const int* exclusionOffsets;
    const int* exclusionPartners;
    int method;
    BoxD box;
    double cutoff;
    PairSettings settings;
    int useSwitch;
    double switchDistance, invSwitchWidth;
    double softcoreR6;          // R^6, 0 without soft core
    int count;                  // lambda values
    double lambdas[MAX_LAMBDA_WINDOWS+1];
};

constexpr int LAMBDA_CHUNK = 256;   // partners scanned by one wave

// One wave per (alchemical particle a, chunk of LAMBDA_CHUNK partners): the energy of the pairs of a with the
// non-alchemical particles of the chunk (the only pairs whose energy depends on lambda: ReferenceInteraxPairIxn.h:100-125
// scales a pair with exactly one alchemical particle by lambda and evaluates it at r_eff), for every lambda, in Real
// arithmetic with FP64 sums.  Membership is the reference's FP64 cutoff decision.  The members of the chunk are first
// compacted into LDS so that every lane of an evaluation batch holds a pair.  out[wave*count+k]: the wave's sum at
// lambda k.
template <typename Real, int TERMS>
__global__ void __launch_bounds__(64) k_lambda_energies(LambdaArgs a, int chunks, double* out) {
    __shared__ int sIndex[LAMBDA_CHUNK];
    __shared__ double sDist[LAMBDA_CHUNK];
    const int i = a.alchList[blockIdx.x/chunks];
    const int first = (blockIdx.x%chunks)*LAMBDA_CHUNK;
    const int lane = threadIdx.x;
    const int origI = (a.atomIndex != nullptr ? a.atomIndex[i] : i);
    double xi, yi, zi;
    load_position(a.posq, a.correction, a.precision, i, xi, yi, zi);
    int members = 0;
    for (int j = first+lane; j < first+LAMBDA_CHUNK; j += 64) {
        bool in = false;
        double r2 = 0;
        if (j < a.n && j != i && (a.flags[j] & ALCHEMICAL_BIT) == 0 &&
            !excluded(a.exclusionOffsets, a.exclusionPartners, i, j)) {
            double xj, yj, zj;
            load_position(a.posq, a.correction, a.precision, j, xj, yj, zj);
            double dx = xj-xi, dy = yj-yi, dz = zj-zi;
            if (a.method == IXH_CUTOFF_PERIODIC)
                reference_min_image(dx, dy, dz, a.box);
            r2 = dx*dx+dy*dy+dz*dz;
            in = (a.method == IXH_NO_CUTOFF ||
                  exact_within_cutoff(a.posq, a.correction, a.precision, a.method == IXH_CUTOFF_PERIODIC, a.box, a.cutoff, i, j));
        }
        const unsigned long long ballot = __ballot(in);
        if (in) {
            const int slot = members+__popcll(ballot & ((1ULL << lane)-1));
            sIndex[slot] = j;
            sDist[slot] = sqrt(r2);
        }
        members += __popcll(ballot);
    }
    __syncthreads();
    PairParticle<Real> P;
    {
        const double* q = a.params+12*(size_t) i;
        P.m1 = Real(q[0]); P.m2 = Real(q[1]); P.m3 = Real(q[2]); P.alpha = Real(q[3]); P.wd = Real(q[4]); P.amp = Real(q[5]);
        P.wx = Real(q[6]); P.iwx = Real(q[7]); P.k = Real(q[8]); P.l = Real(q[9]); P.ws = Real(q[10]); P.iws = Real(q[11]);
    }
    double acc[MAX_LAMBDA_WINDOWS+1];
#pragma unroll
    for (int k = 0; k <= MAX_LAMBDA_WINDOWS; k++)
        acc[k] = 0;
    for (int m = lane; m < members; m += 64) {
        const int j = sIndex[m];
        const Real r = Real(sDist[m]);
        PairParticle<Real> Q;
        {
            const double* q = a.params+12*(size_t) j;
            Q.m1 = Real(q[0]); Q.m2 = Real(q[1]); Q.m3 = Real(q[2]); Q.alpha = Real(q[3]); Q.wd = Real(q[4]); Q.amp = Real(q[5]);
            Q.wx = Real(q[6]); Q.iwx = Real(q[7]); Q.k = Real(q[8]); Q.l = Real(q[9]); Q.ws = Real(q[10]); Q.iws = Real(q[11]);
        }
        Real sw = 1;
        if (a.useSwitch && r > Real(a.switchDistance)) {
            const Real x = (r-Real(a.switchDistance))*Real(a.invSwitchWidth);
            sw = 1+x*x*x*(Real(-10)+x*(Real(15)-x*Real(6)));
        }
        const bool iFirst = origI < (a.atomIndex != nullptr ? a.atomIndex[j] : j);
        Real eHard = 0, dEdR;
        if (a.softcoreR6 == 0.0)
            pair_energy<Real, TERMS>(r, Real(1)/r, P, Q, iFirst, a.settings, eHard, dEdR);
#pragma unroll
        for (int k = 0; k <= MAX_LAMBDA_WINDOWS; k++) {
            if (k >= a.count)
                break;
            const Real lambda = Real(a.lambdas[k]);
            Real e = eHard;
            if (a.softcoreR6 != 0.0) {
                const Real rEff = pow(r*r*r*r*r*r+(Real(1)-lambda)*Real(a.softcoreR6), Real(1.0/6.0));
                pair_energy<Real, TERMS>(rEff, Real(1)/rEff, P, Q, iFirst, a.settings, e, dEdR);
            }
            acc[k] += (double) (lambda*(e*sw));
        }
    }
#pragma unroll
    for (int k = 0; k <= MAX_LAMBDA_WINDOWS; k++) {
        if (k >= a.count)
            break;
        const double total = wave_sum(acc[k]);
        if (lane == 0)
            out[(size_t) blockIdx.x*a.count+k] = total;
    }
}

// sums[k] = sum over the rows w of out[w*count+k], in a fixed order (one wave per column k).
__global__ void __launch_bounds__(64) k_lambda_columns(const double* out, int rows, int count, double* sums) {
    const int k = blockIdx.x;
    double v = 0;
    for (int w = threadIdx.x; w < rows; w += 64)
        v += out[(size_t) w*count+k];
    v = wave_sum(v);
    if (threadIdx.x == 0)
        sums[k] = v;
}

__global__ void k_zero_fixed(long long* p, long long n, double* energy) {
    long long i = (long long) blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n)
        p[i] = 0;
    if (i == 0)
        *energy = 0;
}

}  // namespace

// ------------------------------------------------------------------------------------------------------------
// Static scratch of the evaluations that allocate nothing (ixh_pair_execute_noalloc): the per-workgroup energy and
// pair-count partials of a tile launch, STATIC_PARTIALS of each.  Every GPU the library runs on holds its own copy
// until the process ends; host code reaches it only through the address hipGetSymbolAddress gives on that device.
// ------------------------------------------------------------------------------------------------------------

namespace {
constexpr int STATIC_PARTIALS = 65536;
constexpr size_t STATIC_SCRATCH_BYTES = (size_t) STATIC_PARTIALS*(sizeof(double)+sizeof(int));
constexpr int MAX_DEVICES = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];

namespace {
std::mutex scratchMutex;
void* scratchAddress[MAX_DEVICES] = {};

// The device address of the static scratch on `device`, which must be the current device (resolved once per device).
void* static_scratch(int device) {
    if (device < 0 || device >= MAX_DEVICES)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair: device index out of range");
    std::lock_guard<std::mutex> lock(scratchMutex);
    if (scratchAddress[device] == nullptr) {
        void* ptr = nullptr;
        IXH_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (ptr == nullptr)
            throw Error(IXH_ERROR_HIP, "ixh_pair: the static scratch has no device address");
        scratchAddress[device] = ptr;
    }
    return scratchAddress[device];
}
}  // namespace

// ------------------------------------------------------------------------------------------------------------
// The handle
// ------------------------------------------------------------------------------------------------------------

struct ixh_pair {
    int device = 0;
    int n = 0;
    int terms = 0;
    int method = IXH_NO_CUTOFF;
    double cutoff = 1.0;
    int useSwitch = 0;
    double switchDistance = 0;
    double softcoreRadius = 0;
    int precision = IXH_MIXED;
    bool anyAlchemical = false;
    int exchangeMethod = 0, shortRangeMethod = 0;
    std::vector<unsigned char> alchemical;
    std::vector<int> exclusions;
    double dampingScale = 1;
    int useLRC = 0;
    double lrcCoefficient = 0, lrcAlchemical = 0;
    PairSettings settings;
    // Parameter updates.  Every update writes the FP64 per-particle parameters in the reference's layout (rawDisp 5,
    // rawExch 2, rawSr 3 per particle) and runs the same device work: derived arrays, their slot-ordered copies,
    // class check, long range correction.  updateDone marks its end; what the host still has to do (type-table rows,
    // new classes after a mismatch, reading the correction) happens at the next call that needs it.
    DeviceArray<double> rawDisp, rawExch, rawSr;
    hipEvent_t updateDone = nullptr;
    bool updateRecorded = false;
    bool updatePending = false;    // the next evaluation must wait for updateDone
    bool updateNeedsHost = false;  // ... and, before that, check the classes and form the type-table rows
    bool lrcPending = false;       // the coefficients are in staging once updateDone has completed
    DeviceArray<int> mismatch;
    struct Staging {               // pinned
        double lrc[2];
        int mismatch;
        int unused;
        double reps[10*(MAX_DISPERSION_CLASSES+2*MAX_WIDTH_CLASSES)];
    }* staging = nullptr;
    // type tables: the first particle of every dispersion, exchange and short-range class, and their parameters
    DeviceArray<int> reps;
    DeviceArray<double> repParams;
    int numReps = 0, repExch = 0, repSr = 0;
    // long range correction: classes grouped on the device
    DeviceArray<unsigned long long> lrcKeys[2];
    DeviceArray<int> lrcPerm[2], lrcFirst, lrcIndex, lrcStart;
    DeviceArray<long long> lrcCount;
    DeviceArray<double> lrcClasses, lrcPartial, lrcValues;
    DeviceArray<char> lrcTemp;

    int numBlocks = 0, padded = 0;
    bool built = false;
    int numTiles = 0;
    int pairImageTiles = 0;   // tiles whose pairs take their own minimum image
    // parameters in particle order and in slot order
    DeviceArray<char> paramA, paramB, paramC, sortedA, sortedB, sortedC;
    // block frames (local positions and block centres) of two consecutive calls: with a cutoff, a call's list check
    // writes its frame while the tiles of the previous call still read the other one
    DeviceArray<char> localFrame[2];
    DeviceArray<double> centerFrame[2];
    hipEvent_t frameDone[2] = {nullptr, nullptr};   // the end of the last call that used each frame
    hipStream_t frameStream[2] = {nullptr, nullptr};   // the stream of that call
    bool frameUsed[2] = {false, false};
    int frame = 0;
    DeviceArray<int> tileImage;
    DeviceArray<int> flags, sortedFlags, perm, slotOf, exclusionOffsets, exclusionPartners, blockStart, blockCount, blockList,
        partnerCount, partnerStart, partners, waveCount, tileSlots, tileCandidates, listOffsets;
    DeviceArray<unsigned short> lists;
    long long numCandidates = 0;
    // independent systems (segments) of a batch: particles [segOffsets[s], segOffsets[s+1]), blocks from segBlocks[s],
    // tiles [segTiles[s], segTiles[s+1]); pairs never cross segments
    int numSegments = 0;
    std::vector<int> segOffsets, segBlocks, segTiles;
    DeviceArray<int> segTilesDev;
    DeviceArray<double> tileEnergy;
    int gridLimit = 1 << 20;   // largest number of workgroups of a tile launch (default: one tile per wave)
    // subsets of segments: the tile list and segment list of the last call (pinned staging, device copies)
    StagingRing subsetStaging;
    DeviceArray<int> subsetTiles, subsetSegments;
    // launch order of the tiles: decreasing candidate pairs (cutoff lists) or particle pairs (all-pairs tiles), so
    // that the workgroups dispatched last are the lightest
    DeviceArray<int> tileOrder;
    bool ordered = false;
    // cutoff lists: work items (ranges of the tile array, one per wave, in launch order) of up to tilesPerItem tiles
    int tilesPerItem = 3;
    int numItems = 0;
    DeviceArray<int2> items;
    // type tables (FP32 modes, few parameter classes): classes per particle and the rows of LDS tables
    bool useTables = false;
    int numDisp = 1, numExch = 1, numSr = 1, exchOffset = 0, srOffset = 0, tableFloats = 0;
    int highOrder = 0;
    int singleBranch = 0;   // no width pair of the tables mixes the Taylor and closed-form branches
    bool fixedForm = false;   // the tables allow the fixed table form of the cutoff tile kernel (pair_energy_fixed)
    DeviceArray<int> classes, sortedClasses;
    DeviceArray<float> tables;
    DeviceArray<unsigned> keys, keysOut;
    DeviceArray<int> values;
    DeviceArray<double> pos3, buildPos, bounds, waveEnergy, bbox;
    DeviceArray<int2> tiles;
    DeviceArray<unsigned long long> masks;
    DeviceArray<char> sortTemp;
    DeviceArray<long long> pairCount;
    // energies at several lambda values: FP64 parameters in particle order (12 per particle, written by every
    // parameter update), the alchemical particles, and the scratch buffers of the handle's own evaluation
    DeviceArray<double> lambdaParams;
    DeviceArray<int> alchList;
    int numAlchemical = 0;
    DeviceArray<double> lambdaOut, lambdaEnergy;
    DeviceArray<long long> lambdaForce;
    double* lambdaHost = nullptr;   // pinned
    size_t lambdaHostCount = 0;
    // list-check flags of the two frames, in coherent pinned host memory written by the check kernel itself
    int* hostFlag = nullptr;
    int checkedFrame = 0;              // the frame of the last list check
    hipEvent_t flagCopied = nullptr;   // deferred list checks: the check has finished and its flag is in hostFlag
    bool deferredCheck = false;
    hipStream_t lastStream = nullptr;
    KernelTimer timer;
};

namespace {

void check_desc(const ixh_pair_desc* d) {
    if (d == nullptr)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair: null descriptor");
    if (d->num_particles < 0 || d->num_particles > INDEX_MASK)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair: illegal number of particles");
    if (d->terms <= 0 || d->terms > 7)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair: no pair terms selected");
    if (d->precision != IXH_SINGLE && d->precision != IXH_MIXED && d->precision != IXH_DOUBLE)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair: illegal precision");
    // InteraxPairForce.cpp:56-60, 97-101, 113-136
    if (d->method < 0 || d->method > 2)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "InteraxPairForce: Illegal value for nonbonded method");
    if (d->softcore_radius < 0)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "InteraxPairForce: The soft-core radius cannot be negative");
    for (int e = 0; e < d->num_exclusions; e++) {
        int p1 = d->exclusions[2*e], p2 = d->exclusions[2*e+1];
        if (p1 < 0 || p1 >= d->num_particles || p2 < 0 || p2 >= d->num_particles)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "InteraxPairForce: Illegal particle index for an exclusion");
    }
    if (d->method != IXH_NO_CUTOFF) {
        if (d->cutoff <= 0)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "InteraxPairForce: The cutoff distance must be positive");
        if (d->use_switching_function && (d->switching_distance < 0 || d->switching_distance >= d->cutoff))
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "InteraxPairForce: Switching distance must satisfy 0 <= r_switch < r_cutoff");
        if (d->method == IXH_CUTOFF_PERIODIC && d->default_box != nullptr) {
            const double* b = d->default_box;
            if (d->cutoff > 0.5*b[0] || d->cutoff > 0.5*b[4] || d->cutoff > 0.5*b[8])
                throw Error(IXH_ERROR_INVALID_ARGUMENT,
                            "InteraxPairForce: The cutoff distance cannot be greater than half the periodic box size.");
        }
    }
    auto checkMethod = [](int m) {
        if (m != IXH_OVERLAP_EXACT && m != IXH_OVERLAP_INTERAX_ORDER2)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "InteraxExchangeForce: Illegal value for overlap method");
    };
    if (d->terms & IXH_TERM_DISPERSION) {
        if (d->dispersion == nullptr && d->num_particles > 0)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair: missing dispersion parameters");
    }
    if (d->terms & IXH_TERM_EXCHANGE) {
        if (d->exchange == nullptr && d->num_particles > 0)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair: missing exchange parameters");
        checkMethod(d->exchange_overlap_method);
    }
    if (d->terms & IXH_TERM_SHORT_RANGE) {
        if (d->short_range == nullptr && d->num_particles > 0)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair: missing short range parameters");
        checkMethod(d->short_range_overlap_method);
    }
}

OverlapScheme make_scheme(int method, bool doublePrecision) {
    // InteraxPairForceImpl.h:155-160 with InteraxPairTerms.h:139-159
    OverlapScheme s;
    if (method == IXH_OVERLAP_EXACT) {
        s.order = 6;
        s.about_smaller = 1;
        s.max_delta = 0.0;
        s.max_t = (doublePrecision ? 0.02 : 0.2);
        s.max_tx = (doublePrecision ? 0.3 : 1.6);
    }
    else {
        s.order = 2;
        s.about_smaller = 0;
        s.max_delta = 0.01*0.0529177210903;
        s.max_t = 0.0;
        s.max_tx = HUGE_VAL;
    }
    s.max_delta_f = (float) s.max_delta;
    s.max_t_f = (float) s.max_t;
    s.max_tx_f = (float) s.max_tx;
    return s;
}

// The two rows of an ordered width pair (a, b) of an overlap table, in FP64 rounded once to FP32: the Taylor
// representation, then the closed form.  The branch rule is that of slater_overlap: Taylor where |delta| <
// maxDelta, or |t| < maxT and |t| x < maxTX, written here as r |t|/l < maxTX.
void overlap_rows(double a, double b, const OverlapScheme& s, OverlapRow rows[2]) {
    std::memset(rows, 0, 2*sizeof(OverlapRow));
    if (a == 0 || b == 0)
        return;   // gate 0: no such term for the pair
    const double l = (s.about_smaller ? std::min(a, b) : a);
    const double delta = a+b-l-l;
    const double t = delta/l;
    float kTaylor, txLimit;
    bool taylorUsed = true, closedUsed = true;
    if (std::fabs(delta) < s.max_delta) {
        kTaylor = 0;
        txLimit = 1;
        closedUsed = false;
    }
    else if (std::fabs(t) < s.max_t) {
        kTaylor = (float) (std::fabs(t)/l);
        txLimit = (float) s.max_tx;
    }
    else {
        kTaylor = 0;
        txLimit = 0;
        taylorUsed = false;
    }
    OverlapRow taylor, closed;
    std::memset(&taylor, 0, sizeof(taylor));
    std::memset(&closed, 0, sizeof(closed));
    // exp(-r/l)/(pi l^3) sum_k c_k(t) (r/l)^k
    taylor.k1 = taylor.k2 = (float) (LOG2E/l);
    taylor.n1 = taylor.n2 = (float) (1/l);
    double scale = INV_PI/(l*l*l);
    for (int k = 0; k <= s.order+2; k++, scale /= l)
        taylor.set_coefficient(k, (float) (scale*OverlapTaylor::coefficient(s.order, k, t)));
    const double D = b*b-a*a;
    if (D != 0) {
        const double m1 = 4*a*a*b*b/(D*D*D)*INV_8PI;
        closed.k1 = (float) (LOG2E/a);
        closed.n1 = (float) (1/a);
        closed.k2 = (float) (LOG2E/b);
        closed.n2 = (float) (1/b);
        closed.m1 = (float) m1;
        closed.set_coefficient(0, (float) (a*INV_8PI/(D*D)));
        closed.q1 = (float) -m1;
        closed.q0 = (float) (b*INV_8PI/(D*D));
    }
    else
        closedUsed = false;   // equal widths are always expanded
    if (!taylorUsed)
        taylor = closed;
    if (!closedUsed)
        closed = taylor;
    rows[0] = taylor;
    rows[1] = closed;
    for (int k = 0; k < 2; k++) {
        rows[k].kTaylor = kTaylor;
        rows[k].txLimit = txLimit;
        rows[k].norm = (float) (EIGHT_PI*(a+b)*(a+b)*(a+b));
        rows[k].gate = 1;
        rows[k].high = 0;
        for (int j = 4; j < 9; j++)
            if (rows[k].coefficient(j) != 0.0f)
                rows[k].high = 1;
    }
}

// Whether one overlap serves both overlap terms: both present, the same scheme, and equal widths everywhere.
bool shared_overlap(const ixh_pair* p, const double* exch, const double* sr) {
    bool shared = (p->terms & IXH_TERM_EXCHANGE) && (p->terms & IXH_TERM_SHORT_RANGE) && p->exchangeMethod == p->shortRangeMethod;
    for (int i = 0; shared && i < p->n; i++)
        shared = (exch[2*(size_t) i+1] == sr[3*(size_t) i+2]);
    return shared;
}

// Parameter classes of the type tables (FP32 modes): particles of identical dispersion (m1, m2, m3, alpha), of
// identical exchange (amplitude, width), of identical short range (K, L, width) parameters, or with one overlap of
// identical (amplitude, width, K, L).  Fills cls (dispersion | exchange << 8 | short range << 16 class per particle)
// and reps (the first particle of every dispersion, exchange and short-range class) and returns true; returns false,
// leaving the per-pair path in use, when the particles fall into more classes than the LDS tables hold.
bool form_classes(ixh_pair* p, const double* disp, const double* exch, const double* sr, std::vector<int>& cls, std::vector<int>& reps) {
    const int n = p->n;
    typedef std::tuple<double, double, double, double> Key;
    std::map<Key, int> dispMap, exchMap, srMap;
    std::vector<int> dispReps, exchReps, srReps;
    cls.assign(std::max(n, 1), 0);
    auto classOf = [](std::map<Key, int>& m, std::vector<int>& first, const Key& k, int i) {
        auto it = m.find(k);
        if (it != m.end())
            return it->second;
        const int c = (int) first.size();
        m.emplace(k, c);
        first.push_back(i);
        return c;
    };
    const bool shared = p->settings.shared_overlap != 0;
    const bool ex = (p->terms & IXH_TERM_EXCHANGE) != 0, srTerm = (p->terms & IXH_TERM_SHORT_RANGE) != 0;
    for (int i = 0; i < n; i++) {
        int d = 0, x = 0, s = 0;
        if (p->terms & IXH_TERM_DISPERSION) {
            const double* q = &disp[5*(size_t) i];
            d = classOf(dispMap, dispReps, std::make_tuple(q[0], q[1], q[2], q[3]), i);
        }
        if (shared)
            x = s = classOf(exchMap, exchReps, std::make_tuple(exch[2*(size_t) i+1], exch[2*(size_t) i], sr[3*(size_t) i],
                                                               sr[3*(size_t) i+1]), i);
        else {
            if (ex)
                x = classOf(exchMap, exchReps, std::make_tuple(exch[2*(size_t) i+1], exch[2*(size_t) i], 0.0, 0.0), i);
            if (srTerm)
                s = classOf(srMap, srReps, std::make_tuple(sr[3*(size_t) i+2], 0.0, sr[3*(size_t) i], sr[3*(size_t) i+1]), i);
        }
        if (dispReps.size() > (size_t) MAX_DISPERSION_CLASSES || exchReps.size() > (size_t) MAX_WIDTH_CLASSES ||
            srReps.size() > (size_t) MAX_WIDTH_CLASSES)
            return false;
        cls[i] = d | (x << 8) | (s << 16);
    }
    for (std::vector<int>* r : {&dispReps, &exchReps, &srReps})
        if (r->empty())
            r->push_back(0);
    p->numDisp = (int) dispReps.size();
    p->numExch = (int) exchReps.size();
    p->numSr = (shared ? p->numExch : (int) srReps.size());
    p->exchOffset = 4*p->numDisp*p->numDisp;
    const int exchFloats = OVERLAP_PAIR_FLOATS*p->numExch*p->numExch;
    p->srOffset = (shared ? p->exchOffset : p->exchOffset+exchFloats);
    p->tableFloats = p->srOffset+(shared ? exchFloats : OVERLAP_PAIR_FLOATS*p->numSr*p->numSr);
    reps = dispReps;
    reps.insert(reps.end(), exchReps.begin(), exchReps.end());
    p->repExch = p->numDisp;
    p->repSr = p->repExch;
    if (!shared) {
        p->repSr = p->repExch+p->numExch;
        reps.insert(reps.end(), srReps.begin(), srReps.end());
    }
    p->numReps = (int) reps.size();
    return true;
}

// The rows of the type tables from the parameters of the class representatives (10 per representative: dispersion
// m1, m2, m3, alpha, width; exchange amplitude, width; short range K, L, width).  Also sets highOrder and singleBranch.
void form_tables(ixh_pair* p, const double* rec, std::vector<float>& tables) {
    p->highOrder = 0;
    bool degree3 = false, allDispersion = true;
    tables.assign(p->tableFloats, 0.0f);
    // dispersion rows: C6, C8, C10 combined as in InteraxPairTerms.h:271-279, and the gate alpha_i alpha_j != 0
    if (p->terms & IXH_TERM_DISPERSION)
        for (int ci = 0; ci < p->numDisp; ci++)
            for (int cj = 0; cj < p->numDisp; cj++) {
                const double* qi = rec+10*ci;
                const double* qj = rec+10*cj;
                const double m1i = qi[0], m2i = qi[1], m3i = qi[2], ai = qi[3];
                const double m1j = qj[0], m2j = qj[1], m3j = qj[2], aj = qj[3];
                const double pre = ai*aj/std::max(ai*m1j+aj*m1i, 1e-30);
                float* row = &tables[4*(ci*p->numDisp+cj)];
                row[0] = (float) (pre*m1i*m1j);
                row[1] = (float) (1.5*pre*(m1i*m2j+m2i*m1j));
                row[2] = (float) (pre*(2*(m1i*m3j+m3i*m1j)+4.2*m2i*m2j));
                row[3] = (ai*aj != 0 ? 1.0f : 0.0f);
                if (ai*aj == 0)
                    allDispersion = false;
            }
    // overlap rows of a class pair, with the amplitude product (exchange) and K, L averages (short range); a class
    // key is (width, amplitude, K, L)
    const bool shared = p->settings.shared_overlap != 0;
    auto key = [&](int rep, bool exchange, bool shortRange) {
        const double* q = rec+10*rep;
        return std::make_tuple(exchange ? q[6] : q[9], exchange ? q[5] : 0.0, shortRange ? q[7] : 0.0, shortRange ? q[8] : 0.0);
    };
    auto fill = [&](int firstRep, int count, bool exchange, bool shortRange, int offset, const OverlapScheme& scheme) {
        for (int ci = 0; ci < count; ci++)
            for (int cj = 0; cj < count; cj++) {
                const auto ki = key(firstRep+ci, exchange, shortRange), kj = key(firstRep+cj, exchange, shortRange);
                OverlapRow w[2];
                overlap_rows(std::get<0>(ki), std::get<0>(kj), scheme, w);
                // the energy of the row's terms as S (a1 + a2 S): A_i A_j S + K S_t + L S_t^2 with S_t = norm S
                const double wa = std::get<0>(ki), wb = std::get<0>(kj);
                const double amplitude = std::get<1>(ki)*std::get<1>(kj);
                const double kk = 0.5*(std::get<2>(ki)+std::get<2>(kj)), ll = 0.5*(std::get<3>(ki)+std::get<3>(kj));
                const double norm = (wa == 0 || wb == 0 ? 0.0 : EIGHT_PI*(wa+wb)*(wa+wb)*(wa+wb));
                for (OverlapRow& row : w) {
                    if (row.high != 0.0f)
                        p->highOrder = 1;
                    if (row.p3 != 0.0f)
                        degree3 = true;
                    row.amplitude = (float) amplitude;
                    row.kk = (float) kk;
                    row.ll = (float) ll;
                    row.a1 = (float) (amplitude+kk*norm);
                    row.a2 = (float) (ll*norm*norm);
                }
                std::memcpy(&tables[offset+OVERLAP_PAIR_FLOATS*(ci*count+cj)], w, sizeof(w));
            }
    };
    if (shared)
        fill(p->repExch, p->numExch, true, true, p->exchOffset, p->settings.exchange_scheme);
    else {
        if (p->terms & IXH_TERM_EXCHANGE)
            fill(p->repExch, p->numExch, true, false, p->exchOffset, p->settings.exchange_scheme);
        if (p->terms & IXH_TERM_SHORT_RANGE)
            fill(p->repSr, p->numSr, false, true, p->srOffset, p->settings.short_range_scheme);
    }
    // rows of a width pair that takes one branch only are equal: the kernel then reads the first without selecting
    p->singleBranch = 1;
    for (int offset = p->exchOffset; offset < p->tableFloats; offset += OVERLAP_PAIR_FLOATS)
        if (std::memcmp(&tables[offset], &tables[offset+OVERLAP_ROW_FLOATS], sizeof(OverlapRow)) != 0)
            p->singleBranch = 0;
    const bool bothOverlaps = (p->terms & IXH_TERM_EXCHANGE) && (p->terms & IXH_TERM_SHORT_RANGE);
    p->fixedForm = p->singleBranch && !p->highOrder && !degree3 && (allDispersion || !(p->terms & IXH_TERM_DISPERSION)) &&
                   (shared || !bothOverlaps);
}

// Classes, shared-overlap choice and type-table layout from host copies of the parameters (at creation, and when an
// update no longer fits the classes).  Synchronous.
void classify(ixh_pair* p, const double* disp, const double* exch, const double* sr) {
    p->settings.shared_overlap = shared_overlap(p, exch, sr) ? 1 : 0;
    p->useTables = false;
    p->numReps = 0;
    if (p->precision == IXH_DOUBLE)
        return;
    std::vector<int> cls, reps;
    p->useTables = form_classes(p, disp, exch, sr, cls, reps);
    if (p->useTables) {
        p->classes.reserve(cls.size());
        IXH_HIP_CHECK(hipMemcpy(p->classes.ptr, cls.data(), sizeof(int)*cls.size(), hipMemcpyHostToDevice));
        p->reps.reserve(reps.size());
        IXH_HIP_CHECK(hipMemcpy(p->reps.ptr, reps.data(), sizeof(int)*reps.size(), hipMemcpyHostToDevice));
    }
}

// The slot-ordered copies of the per-particle arrays (after a list build, or a parameter update of a built list).
void gather_sorted(ixh_pair* p, hipStream_t stream) {
    if (p->precision == IXH_DOUBLE)
        hipLaunchKernelGGL(k_gather_params<double>, dim3(grid_for(p->padded, 256)), dim3(256), 0, stream, p->perm.ptr, p->padded,
                           (const double4*) p->paramA.ptr, (const double4*) p->paramB.ptr, (const double4*) p->paramC.ptr, p->flags.ptr,
                           (double4*) p->sortedA.ptr, (double4*) p->sortedB.ptr, (double4*) p->sortedC.ptr, p->sortedFlags.ptr);
    else
        hipLaunchKernelGGL(k_gather_params<float>, dim3(grid_for(p->padded, 256)), dim3(256), 0, stream, p->perm.ptr, p->padded,
                           (const float4*) p->paramA.ptr, (const float4*) p->paramB.ptr, (const float4*) p->paramC.ptr, p->flags.ptr,
                           (float4*) p->sortedA.ptr, (float4*) p->sortedB.ptr, (float4*) p->sortedC.ptr, p->sortedFlags.ptr);
    IXH_HIP_CHECK(hipGetLastError());
    if (p->useTables) {
        p->sortedClasses.reserve(p->padded);
        hipLaunchKernelGGL(k_gather_int, dim3(grid_for(p->padded, 256)), dim3(256), 0, stream, p->perm.ptr, p->padded,
                           (const int*) p->classes.ptr, p->sortedClasses.ptr);
        IXH_HIP_CHECK(hipGetLastError());
    }
}

// The evaluation's arrays from the FP64 parameters, the class check, the representatives' parameters (type tables),
// and the slot-ordered copies when a list is built.
void derive_parameters(ixh_pair* p, hipStream_t stream) {
    const int n = p->n;
    const size_t r4 = (p->precision == IXH_DOUBLE ? sizeof(double4) : sizeof(float4));
    p->paramA.reserve(r4*std::max(n, 1));
    p->paramB.reserve(r4*std::max(n, 1));
    p->paramC.reserve(r4*std::max(n, 1));
    p->lambdaParams.reserve(12*(size_t) std::max(n, 1));
    hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(1), 0, stream, p->mismatch.ptr);
    IXH_HIP_CHECK(hipGetLastError());
    DeriveArgs a;
    a.disp = (p->terms & IXH_TERM_DISPERSION) ? p->rawDisp.ptr : nullptr;
    a.exch = (p->terms & IXH_TERM_EXCHANGE) ? p->rawExch.ptr : nullptr;
    a.sr = (p->terms & IXH_TERM_SHORT_RANGE) ? p->rawSr.ptr : nullptr;
    a.n = n;
    a.precision = p->precision;
    a.useTables = p->useTables ? 1 : 0;
    a.shared = p->settings.shared_overlap;
    a.cls = p->classes.ptr;
    a.reps = p->reps.ptr;
    a.repExch = p->repExch;
    a.repSr = p->repSr;
    a.A = p->paramA.ptr;
    a.B = p->paramB.ptr;
    a.C = p->paramC.ptr;
    a.lambdaParams = p->lambdaParams.ptr;
    a.mismatch = p->mismatch.ptr;
    if (n > 0) {
        hipLaunchKernelGGL(k_derive_params, dim3(grid_for(n, 256)), dim3(256), 0, stream, a);
        IXH_HIP_CHECK(hipGetLastError());
    }
    IXH_HIP_CHECK(hipMemcpyAsync(&p->staging->mismatch, p->mismatch.ptr, sizeof(int), hipMemcpyDeviceToHost, stream));
    if (p->useTables) {
        p->repParams.reserve(10*(size_t) p->numReps);
        hipLaunchKernelGGL(k_gather_reps, dim3(1), dim3(64), 0, stream, a.disp, a.exch, a.sr, (const int*) p->reps.ptr, p->numReps,
                           p->repParams.ptr);
        IXH_HIP_CHECK(hipGetLastError());
        IXH_HIP_CHECK(hipMemcpyAsync(p->staging->reps, p->repParams.ptr, sizeof(double)*10*p->numReps, hipMemcpyDeviceToHost, stream));
    }
    if (p->built)
        gather_sorted(p, stream);
}

// Long range correction coefficients (InteraxDispersionForceImpl.cpp:86-149) on the device, into the staging area:
// classes of identical (m1, m2, m3, alpha, width, alchemical) in the order of the tuple, grouped by stable radix
// sorts; the class-pair integrals of k_long_range; their sum in a fixed order.
void launch_long_range(ixh_pair* p, hipStream_t stream) {
    const int n = p->n;
    if (!(p->terms & IXH_TERM_DISPERSION) || p->method != IXH_CUTOFF_PERIODIC || !p->useLRC || n == 0) {
        p->lrcCoefficient = p->lrcAlchemical = 0;
        p->lrcPending = false;
        return;
    }
    for (int b = 0; b < 2; b++) {
        p->lrcKeys[b].reserve(n);
        p->lrcPerm[b].reserve(n);
    }
    p->lrcFirst.reserve(n);
    p->lrcIndex.reserve(n);
    p->lrcStart.reserve(n+1);
    p->lrcCount.reserve(1);
    p->lrcClasses.reserve(7*(size_t) n);
    p->lrcPartial.reserve(2*LRC_BLOCKS);
    p->lrcValues.reserve(2);
    size_t sortBytes = 0, scanBytes = 0;
    IXH_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, sortBytes, p->lrcKeys[0].ptr, p->lrcKeys[1].ptr, p->lrcPerm[0].ptr,
                                            p->lrcPerm[1].ptr, (size_t) n, 0, 64, stream));
    IXH_HIP_CHECK(rocprim::inclusive_scan(nullptr, scanBytes, p->lrcFirst.ptr, p->lrcIndex.ptr, (size_t) n, rocprim::plus<int>(),
                                          stream));
    size_t tempBytes = std::max(sortBytes, scanBytes);
    p->lrcTemp.reserve(tempBytes);
    const int grid = grid_for(n, 256);
    const int* flags = p->flags.ptr;
    hipLaunchKernelGGL(k_iota, dim3(grid), dim3(256), 0, stream, p->lrcPerm[0].ptr, n);
    IXH_HIP_CHECK(hipGetLastError());
    // least significant field first: alchemical, width, alpha, m3, m2, m1
    for (int field = 5; field >= 0; field--) {
        hipLaunchKernelGGL(k_lrc_keys, dim3(grid), dim3(256), 0, stream, p->rawDisp.ptr, flags, n, field, p->lrcPerm[0].ptr,
                           p->lrcKeys[0].ptr);
        IXH_HIP_CHECK(hipGetLastError());
        size_t bytes = tempBytes;
        IXH_HIP_CHECK(rocprim::radix_sort_pairs(p->lrcTemp.ptr, bytes, p->lrcKeys[0].ptr, p->lrcKeys[1].ptr, p->lrcPerm[0].ptr,
                                                p->lrcPerm[1].ptr, (size_t) n, 0, 64, stream));
        std::swap(p->lrcPerm[0].ptr, p->lrcPerm[1].ptr);   // both hold n entries
        std::swap(p->lrcPerm[0].count, p->lrcPerm[1].count);
        std::swap(p->lrcPerm[0].bytes, p->lrcPerm[1].bytes);
    }
    hipLaunchKernelGGL(k_lrc_bounds, dim3(grid), dim3(256), 0, stream, p->rawDisp.ptr, flags, p->lrcPerm[0].ptr, n, p->lrcFirst.ptr);
    IXH_HIP_CHECK(hipGetLastError());
    size_t bytes = tempBytes;
    IXH_HIP_CHECK(rocprim::inclusive_scan(p->lrcTemp.ptr, bytes, p->lrcFirst.ptr, p->lrcIndex.ptr, (size_t) n, rocprim::plus<int>(),
                                          stream));
    hipLaunchKernelGGL(k_lrc_starts, dim3(grid), dim3(256), 0, stream, p->lrcFirst.ptr, p->lrcIndex.ptr, n, p->lrcStart.ptr,
                       p->lrcCount.ptr);
    IXH_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_lrc_classes, dim3(grid), dim3(256), 0, stream, p->rawDisp.ptr, flags, p->lrcPerm[0].ptr, p->lrcStart.ptr,
                       p->lrcCount.ptr, p->lrcClasses.ptr);
    IXH_HIP_CHECK(hipGetLastError());
    LrcArgs a;
    a.classes = p->lrcClasses.ptr;
    a.numClasses = p->lrcCount.ptr;
    a.cutoff = p->cutoff;
    a.useSwitch = p->useSwitch;
    a.switchDistance = (p->useSwitch ? p->switchDistance : p->cutoff);
    a.switchWidth = p->cutoff-a.switchDistance;
    a.dampingScale = p->dampingScale;
    gauss_legendre(32, a.nodes, a.weights);
    hipLaunchKernelGGL(k_long_range, dim3(LRC_BLOCKS), dim3(LRC_THREADS), 0, stream, a, p->lrcPartial.ptr);
    IXH_HIP_CHECK(hipGetLastError());
    const double N = n;
    const double normalization = 2*M_PI*N*N/(0.5*N*(N+1));
    hipLaunchKernelGGL(k_lrc_reduce, dim3(1), dim3(1024), 0, stream, (const double*) p->lrcPartial.ptr, LRC_BLOCKS, normalization,
                       p->lrcValues.ptr);
    IXH_HIP_CHECK(hipGetLastError());
    IXH_HIP_CHECK(hipMemcpyAsync(p->staging->lrc, p->lrcValues.ptr, 2*sizeof(double), hipMemcpyDeviceToHost, stream));
    p->lrcPending = true;
}

// The device work of a parameter update, once the FP64 parameters are in rawDisp, rawExch, rawSr (stream order).
void update_parameters(ixh_pair* p, double dampingScale, hipStream_t stream) {
    p->dampingScale = dampingScale;
    p->settings.damping_scale = dampingScale;
    p->settings.damping_scale_f = (float) dampingScale;
    derive_parameters(p, stream);
    launch_long_range(p, stream);
    IXH_HIP_CHECK(hipEventRecord(p->updateDone, stream));
    p->updateRecorded = true;
    p->updatePending = true;
    p->updateNeedsHost = (p->useTables || p->settings.shared_overlap);
}

// Before an update writes the handle's arrays: the evaluations and the update before it have finished with them.
void order_update(ixh_pair* p, hipStream_t stream) {
    for (int f = 0; f < 2; f++)
        if (p->frameStream[f] != stream && p->frameUsed[f])
            IXH_HIP_CHECK(hipStreamWaitEvent(stream, p->frameDone[f], 0));
    if (p->updateRecorded)
        IXH_HIP_CHECK(hipStreamWaitEvent(stream, p->updateDone, 0));
}

// The host's part of the last update, before an evaluation on `stream` (which then waits for the update's device
// work).  With type tables (or one shared overlap) it waits for the update: the class check decides whether the
// classes still hold (then the table rows are formed from the representatives' parameters) or are formed anew.
void finish_parameters(ixh_pair* p, hipStream_t stream) {
    if (!p->updatePending)
        return;
    if (p->updateNeedsHost) {
        IXH_HIP_CHECK(hipEventSynchronize(p->updateDone));
        if (p->staging->mismatch) {
            const int n = p->n;
            std::vector<double> disp(5*(size_t) n), exch(2*(size_t) n), sr(3*(size_t) n);
            if (p->terms & IXH_TERM_DISPERSION)
                IXH_HIP_CHECK(hipMemcpy(disp.data(), p->rawDisp.ptr, sizeof(double)*disp.size(), hipMemcpyDeviceToHost));
            if (p->terms & IXH_TERM_EXCHANGE)
                IXH_HIP_CHECK(hipMemcpy(exch.data(), p->rawExch.ptr, sizeof(double)*exch.size(), hipMemcpyDeviceToHost));
            if (p->terms & IXH_TERM_SHORT_RANGE)
                IXH_HIP_CHECK(hipMemcpy(sr.data(), p->rawSr.ptr, sizeof(double)*sr.size(), hipMemcpyDeviceToHost));
            classify(p, disp.data(), exch.data(), sr.data());
            derive_parameters(p, nullptr);
            IXH_HIP_CHECK(hipStreamSynchronize(nullptr));
            if (p->staging->mismatch)
                throw Error(IXH_ERROR_HIP, "ixh_pair: the parameter classes do not hold after they were formed anew");
        }
        if (p->useTables) {
            std::vector<float> tables;
            form_tables(p, p->staging->reps, tables);
            p->tables.reserve(tables.size());
            IXH_HIP_CHECK(hipMemcpy(p->tables.ptr, tables.data(), sizeof(float)*tables.size(), hipMemcpyHostToDevice));
        }
        p->updateNeedsHost = false;
    }
    IXH_HIP_CHECK(hipStreamWaitEvent(stream, p->updateDone, 0));
    p->updatePending = false;
}

// The long range correction of the last update (waits for it).
void finish_long_range(ixh_pair* p) {
    if (!p->lrcPending)
        return;
    IXH_HIP_CHECK(hipEventSynchronize(p->updateDone));
    p->lrcCoefficient = p->staging->lrc[0];
    p->lrcAlchemical = p->staging->lrc[1];
    p->lrcPending = false;
}

// Checks of an update (ReferenceInteraxKernels.cpp:98-102, 142-148, 185-191).
void check_update(const ixh_pair* p, int terms, int exchangeMethod, int shortRangeMethod) {
    if (terms != p->terms)
        throw Error(IXH_ERROR_STATE, "updateParametersInContext: The set of pair forces has changed");
    if (((p->terms & IXH_TERM_EXCHANGE) && exchangeMethod != p->exchangeMethod) ||
        ((p->terms & IXH_TERM_SHORT_RANGE) && shortRangeMethod != p->shortRangeMethod))
        throw Error(IXH_ERROR_STATE, "updateParametersInContext: The overlap method has changed");
}

// Host arrays of a descriptor into the FP64 device arrays (synchronous: the caller may reuse them on return).
void upload_raw(ixh_pair* p, const ixh_pair_desc* d) {
    const size_t n = std::max(p->n, 1);
    p->rawDisp.reserve(5*n);
    p->rawExch.reserve(2*n);
    p->rawSr.reserve(3*n);
    if (p->n == 0)
        return;
    if (p->terms & IXH_TERM_DISPERSION)
        IXH_HIP_CHECK(hipMemcpy(p->rawDisp.ptr, d->dispersion, sizeof(double)*5*(size_t) p->n, hipMemcpyHostToDevice));
    if (p->terms & IXH_TERM_EXCHANGE)
        IXH_HIP_CHECK(hipMemcpy(p->rawExch.ptr, d->exchange, sizeof(double)*2*(size_t) p->n, hipMemcpyHostToDevice));
    if (p->terms & IXH_TERM_SHORT_RANGE)
        IXH_HIP_CHECK(hipMemcpy(p->rawSr.ptr, d->short_range, sizeof(double)*3*(size_t) p->n, hipMemcpyHostToDevice));
}

// The parameters of a new handle: schemes, particle flags, alchemical list, classes, then the common update.
void init_parameters(ixh_pair* p, const ixh_pair_desc* d, hipStream_t stream) {
    const int n = p->n;
    const bool doublePrecision = (p->precision == IXH_DOUBLE);
    p->settings.exchange_scheme = make_scheme(p->exchangeMethod, doublePrecision);
    p->settings.short_range_scheme = make_scheme(p->shortRangeMethod, doublePrecision);
    std::vector<int> flags(std::max(n, 1), 0), alch;
    for (int i = 0; i < n; i++) {
        flags[i] = i | (p->alchemical[i] ? ALCHEMICAL_BIT : 0);
        if (p->alchemical[i])
            alch.push_back(i);
    }
    p->flags.upload(flags.data(), flags.size(), stream);
    p->alchList.reserve(std::max<size_t>(alch.size(), 1));
    if (!alch.empty())
        p->alchList.upload(alch.data(), alch.size(), stream);
    p->numAlchemical = (int) alch.size();
    IXH_HIP_CHECK(hipStreamSynchronize(stream));
    upload_raw(p, d);
    const double* disp = d->dispersion;
    const double* exch = d->exchange;
    const double* sr = d->short_range;
    classify(p, disp, exch, sr);
    update_parameters(p, (p->terms & IXH_TERM_DISPERSION) ? d->damping_scale : 1.0, stream);
    finish_parameters(p, stream);
    finish_long_range(p);
}

// Build the slot order, the partner lists, the tiles and the candidate masks from the current positions.
void build_list(ixh_pair* p, const ixh_buffers* b, hipStream_t stream) {
    int n = p->n;
    p->numItems = 0;
    bool periodic = (p->method == IXH_CUTOFF_PERIODIC);
    bool cutoff = (p->method != IXH_NO_CUTOFF);
    BoxD box = make_box(b->box);
    const float4* corr = (const float4*) b->posq_correction;
    int Nb = p->numBlocks;
    p->pos3.reserve(3*(size_t) n);
    p->perm.reserve(p->padded);
    p->slotOf.reserve(n);
    std::vector<int2> tiles;
    if (!cutoff) {
        // all pairs: the particle order is kept, every block J > I is a partner chunk; with segments, each segment
        // starts a new block and only blocks of the same segment pair up
        std::vector<int> perm(p->padded, -1);
        if (p->numSegments > 0) {
            p->segTiles.assign(p->numSegments+1, 0);
            for (int g = 0; g < p->numSegments; g++) {
                const int first = p->segOffsets[g], count = p->segOffsets[g+1]-first;
                const int B0 = p->segBlocks[g], B1 = p->segBlocks[g+1];
                for (int k = 0; k < count; k++)
                    perm[B0*BLOCK+k] = first+k;
                for (int I = B0; I < B1; I++) {
                    tiles.push_back(make_int2(I, -1));
                    for (int J = I+1; J < B1; J++)
                        tiles.push_back(make_int2(I, J*BLOCK));
                }
                p->segTiles[g+1] = (int) tiles.size();
            }
            p->segTilesDev.upload(p->segTiles.data(), p->segTiles.size(), stream);
        }
        else {
            for (int s = 0; s < n; s++)
                perm[s] = s;
            for (int I = 0; I < Nb; I++) {
                tiles.push_back(make_int2(I, -1));
                for (int J = I+1; J < Nb; J++)
                    tiles.push_back(make_int2(I, J*BLOCK));
            }
        }
        p->perm.upload(perm.data(), perm.size(), stream);
        std::vector<int> partners(p->padded);
        std::iota(partners.begin(), partners.end(), 0);
        p->partners.upload(partners.data(), partners.size(), stream);
        p->pairImageTiles = 0;
        // launch order: tiles with the most particle pairs first
        std::vector<int> valid(p->padded/BLOCK, 0);
        for (int slot = 0; slot < p->padded; slot++)
            if (perm[slot] >= 0)
                valid[slot/BLOCK]++;
        std::vector<long long> work(tiles.size());
        for (size_t t = 0; t < tiles.size(); t++) {
            const long long vi = valid[tiles[t].x];
            work[t] = (tiles[t].y < 0 ? vi*(vi-1)/2 : vi*valid[tiles[t].y/BLOCK]);
        }
        std::vector<int> order(tiles.size());
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
        p->tileOrder.upload(order.data(), std::max<size_t>(order.size(), 1), stream);
        p->ordered = !tiles.empty();
    }
    else {
        KeyArgs k;
        k.posq = b->posq;
        k.correction = corr;
        k.precision = b->precision;
        k.n = n;
        k.periodic = periodic;
        k.box = box;
        if (!periodic) {
            p->bbox.reserve(6);
            hipLaunchKernelGGL(k_bounding_box, dim3(1), dim3(256), 0, stream, b->posq, corr, b->precision, n, p->bbox.ptr);
            IXH_HIP_CHECK(hipGetLastError());
            double bb[6];
            IXH_HIP_CHECK(hipMemcpyAsync(bb, p->bbox.ptr, sizeof(bb), hipMemcpyDeviceToHost, stream));
            IXH_HIP_CHECK(hipStreamSynchronize(stream));
            for (int c = 0; c < 3; c++) {
                k.lo[c] = bb[c];
                double extent = bb[3+c]-bb[c];
                k.scale[c] = (extent > 0 ? 1.0/(extent*(1+1e-9)) : 0.0);
            }
        }
        p->keys.reserve(n);
        p->keysOut.reserve(n);
        p->values.reserve(n);
        hipLaunchKernelGGL(k_morton_keys, dim3(grid_for(n, 256)), dim3(256), 0, stream, k, p->keys.ptr, p->values.ptr, p->pos3.ptr);
        IXH_HIP_CHECK(hipGetLastError());
        size_t tempBytes = 0;
        IXH_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tempBytes, p->keys.ptr, p->keysOut.ptr, p->values.ptr, p->perm.ptr,
                                                (size_t) n, 0, 30, stream));
        p->sortTemp.reserve(tempBytes);
        IXH_HIP_CHECK(rocprim::radix_sort_pairs(p->sortTemp.ptr, tempBytes, p->keys.ptr, p->keysOut.ptr, p->values.ptr, p->perm.ptr,
                                                (size_t) n, 0, 30, stream));
    }
    if (cutoff) {
        // the sort fills slots [0, n): the rest are padding
        hipLaunchKernelGGL(k_finish_perm, dim3(grid_for(p->padded, 256)), dim3(256), 0, stream, p->perm.ptr, n, p->padded, p->slotOf.ptr);
        IXH_HIP_CHECK(hipGetLastError());
    }
    size_t r4 = (p->precision == IXH_DOUBLE ? sizeof(double4) : sizeof(float4));
    p->sortedA.reserve(r4*p->padded);
    p->sortedB.reserve(r4*p->padded);
    p->sortedC.reserve(r4*p->padded);
    p->sortedFlags.reserve(p->padded);
    gather_sorted(p, stream);

    double listCutoff = p->cutoff+SKIN;
    if (cutoff) {
        // candidate blocks from bounding boxes, then candidate particles against the box of each block
        p->bounds.reserve(6*(size_t) Nb);
        BlockBoundsArgs bb;
        bb.pos3 = p->pos3.ptr;
        bb.perm = p->perm.ptr;
        bb.numBlocks = Nb;
        bb.periodic = periodic;
        bb.box = box;
        hipLaunchKernelGGL(k_block_bounds, dim3(grid_for(Nb, 64)), dim3(64), 0, stream, bb, p->bounds.ptr);
        IXH_HIP_CHECK(hipGetLastError());
        p->blockCount.reserve(Nb);
        hipLaunchKernelGGL(k_find_blocks, dim3(Nb), dim3(64), 0, stream, p->bounds.ptr, Nb, (int) periodic, box, listCutoff*listCutoff,
                           p->blockCount.ptr, (const int*) nullptr, (int*) nullptr);
        IXH_HIP_CHECK(hipGetLastError());
        std::vector<int> counts(Nb), start(Nb+1, 0);
        std::vector<double> bounds(6*(size_t) Nb);
        IXH_HIP_CHECK(hipMemcpyAsync(counts.data(), p->blockCount.ptr, sizeof(int)*Nb, hipMemcpyDeviceToHost, stream));
        IXH_HIP_CHECK(hipMemcpyAsync(bounds.data(), p->bounds.ptr, sizeof(double)*bounds.size(), hipMemcpyDeviceToHost, stream));
        IXH_HIP_CHECK(hipStreamSynchronize(stream));
        for (int I = 0; I < Nb; I++)
            start[I+1] = start[I]+counts[I];
        p->blockStart.upload(start.data(), start.size(), stream);
        p->blockList.reserve(std::max(start[Nb], 1));
        hipLaunchKernelGGL(k_find_blocks, dim3(Nb), dim3(64), 0, stream, p->bounds.ptr, Nb, (int) periodic, box, listCutoff*listCutoff,
                           p->blockCount.ptr, (const int*) p->blockStart.ptr, p->blockList.ptr);
        IXH_HIP_CHECK(hipGetLastError());
        p->partnerCount.reserve(Nb);
        hipLaunchKernelGGL(k_find_partners, dim3(Nb), dim3(64), 0, stream, p->bounds.ptr, p->pos3.ptr, p->perm.ptr, p->blockStart.ptr,
                           p->blockList.ptr, (int) periodic, box, listCutoff*listCutoff, p->partnerCount.ptr, (const int*) nullptr,
                           (int*) nullptr);
        IXH_HIP_CHECK(hipGetLastError());
        IXH_HIP_CHECK(hipMemcpyAsync(counts.data(), p->partnerCount.ptr, sizeof(int)*Nb, hipMemcpyDeviceToHost, stream));
        IXH_HIP_CHECK(hipStreamSynchronize(stream));
        std::vector<int> pstart(Nb+1, 0);
        for (int I = 0; I < Nb; I++)
            pstart[I+1] = pstart[I]+(counts[I]+BLOCK-1)/BLOCK*BLOCK;
        p->partnerStart.upload(pstart.data(), pstart.size(), stream);
        p->partners.reserve(std::max(pstart[Nb], 1));
        IXH_HIP_CHECK(hipMemsetAsync(p->partners.ptr, 0xff, sizeof(int)*(size_t) std::max(pstart[Nb], 1), stream));
        hipLaunchKernelGGL(k_find_partners, dim3(Nb), dim3(64), 0, stream, p->bounds.ptr, p->pos3.ptr, p->perm.ptr, p->blockStart.ptr,
                           p->blockList.ptr, (int) periodic, box, listCutoff*listCutoff, p->partnerCount.ptr,
                           (const int*) p->partnerStart.ptr, p->partners.ptr);
        IXH_HIP_CHECK(hipGetLastError());
        for (int I = 0; I < Nb; I++) {
            tiles.push_back(make_int2(I, -1));
            for (int off = pstart[I]; off < pstart[I+1]; off += BLOCK)
                tiles.push_back(make_int2(I, off));
        }
    }
    p->numTiles = (int) tiles.size();
    p->tiles.upload(tiles.data(), tiles.size(), stream);
    int T = p->numTiles;
    p->masks.reserve((size_t) std::max(T, 1)*BLOCK);
    p->tileSlots.reserve((size_t) std::max(T, 1)*BLOCK);
    MaskArgs m;
    m.tiles = p->tiles.ptr;
    m.partners = p->partners.ptr;
    m.perm = p->perm.ptr;
    m.pos3 = p->pos3.ptr;
    m.exclusionOffsets = p->exclusionOffsets.ptr;
    m.exclusionPartners = p->exclusionPartners.ptr;
    m.useCutoff = cutoff;
    m.periodic = periodic;
    m.box = box;
    m.listCutoff2 = listCutoff*listCutoff;
    m.bounds = p->bounds.ptr;
    // Pairs need their own minimum image unless every pair in range is the nearest image of its block shift: cutoff +
    // the extents of both blocks (each grows by at most the skin before a rebuild) below half the box (per tile)
    m.imageLimit = 0.5*std::min(b->box[0], std::min(b->box[4], b->box[8]))-p->cutoff-2*SKIN;
    p->tileImage.reserve(std::max(T, 1));
    if (T > 0) {
        hipLaunchKernelGGL(k_build_masks, dim3(T), dim3(BLOCK), 0, stream, m, p->masks.ptr, p->tileSlots.ptr, p->tileImage.ptr);
        IXH_HIP_CHECK(hipGetLastError());
    }
    p->numCandidates = 0;
    if (cutoff && T > 0) {
        // candidate lists: count, lay out (each tile padded to a multiple of 4 entries), write
        p->tileCandidates.reserve(T);
        hipLaunchKernelGGL(k_candidate_lists, dim3(T), dim3(BLOCK), 0, stream, p->tiles.ptr, p->masks.ptr, p->tileCandidates.ptr,
                           (const int*) nullptr, (unsigned short*) nullptr);
        IXH_HIP_CHECK(hipGetLastError());
        std::vector<int> unsortedCounts(T), counts(T), offsets(T+1, 0), image(T);
        IXH_HIP_CHECK(hipMemcpyAsync(unsortedCounts.data(), p->tileCandidates.ptr, sizeof(int)*T, hipMemcpyDeviceToHost, stream));
        IXH_HIP_CHECK(hipMemcpyAsync(image.data(), p->tileImage.ptr, sizeof(int)*T, hipMemcpyDeviceToHost, stream));
        IXH_HIP_CHECK(hipStreamSynchronize(stream));
        {
            // Work items: up to tilesPerItem consecutive tiles of one block I (its diagonal tile, then its partner
            // chunks), one item per wave, so that the i side is staged and flushed once per item; items launched
            // heaviest first (the workgroups dispatched last are the lightest), padded with empty items to whole
            // workgroups.  The tiles are stored in launch order, so that a wave reads its tiles with no index in
            // between.  (Items grouped by spatial region per XCD, for L2 locality, measured 10-15 % slower on the
            // 98 304-atom box than this global weight order.)
            std::vector<std::pair<long long, int2>> sortedItems;   // (weight, (first tile, tiles))
            for (int t0 = 0; t0 < T;) {
                const int I = tiles[t0].x;
                int t1 = t0;
                long long weight = 0;
                while (t1 < T && tiles[t1].x == I && t1-t0 < p->tilesPerItem)
                    weight += unsortedCounts[t1++];
                sortedItems.push_back(std::make_pair(weight, make_int2(t0, t1-t0)));
                t0 = t1;
            }
            std::stable_sort(sortedItems.begin(), sortedItems.end(),
                             [](const std::pair<long long, int2>& x, const std::pair<long long, int2>& y) { return x.first > y.first; });
            std::vector<int> order;
            order.reserve(T);
            std::vector<int2> itemRanges;
            const size_t numItems = (sortedItems.size()+QUEUE_GROUP_WAVES-1)/QUEUE_GROUP_WAVES*QUEUE_GROUP_WAVES;
            itemRanges.reserve(numItems);
            for (size_t k = 0; k < numItems; k++) {
                const int begin = (int) order.size();
                if (k < sortedItems.size()) {
                    const int2 item = sortedItems[k].second;
                    for (int t = item.x; t < item.x+item.y; t++)
                        order.push_back(t);
                }
                itemRanges.push_back(make_int2(begin, (int) order.size()));
            }
            if (order.size() != (size_t) T)
                throw Error(IXH_ERROR_HIP, "ixh_pair: work items do not cover the tiles");
            p->numItems = (int) itemRanges.size();
            p->items.upload(itemRanges.data(), std::max<size_t>(itemRanges.size(), 1), stream);
            std::vector<int2> sortedTiles(T);
            p->pairImageTiles = 0;
            for (int t = 0; t < T; t++) {
                sortedTiles[t] = tiles[order[t]];
                counts[t] = unsortedCounts[order[t]];
                if (image[order[t]]) {
                    sortedTiles[t].x |= TILE_PAIR_IMAGE;
                    p->pairImageTiles++;
                }
            }
            p->tileOrder.upload(order.data(), order.size(), stream);
            DeviceArray<unsigned long long> masks;
            DeviceArray<int> slots;
            masks.reserve((size_t) T*BLOCK);
            slots.reserve((size_t) T*BLOCK);
            hipLaunchKernelGGL(k_gather_tile_rows, dim3(T), dim3(BLOCK), 0, stream, (const int*) p->tileOrder.ptr,
                               (const unsigned long long*) p->masks.ptr, (const int*) p->tileSlots.ptr, masks.ptr, slots.ptr);
            IXH_HIP_CHECK(hipGetLastError());
            std::swap(p->masks.ptr, masks.ptr);
            std::swap(p->masks.count, masks.count);
            std::swap(p->tileSlots.ptr, slots.ptr);
            std::swap(p->tileSlots.count, slots.count);
            p->tiles.upload(sortedTiles.data(), sortedTiles.size(), stream);
            IXH_HIP_CHECK(hipStreamSynchronize(stream));   // before the unsorted rows are freed
        }
        p->ordered = false;
        long long total = 0;
        for (int t = 0; t < T; t++) {
            p->numCandidates += counts[t];
            total += (counts[t]+3) & ~3;
            if (total > INT32_MAX)
                throw Error(IXH_ERROR_CAPACITY, "ixh_pair_execute: the candidate pair list exceeds 2^31 entries");
            offsets[t+1] = (int) total;
        }
        p->listOffsets.upload(offsets.data(), offsets.size(), stream);
        p->lists.reserve((size_t) std::max<long long>(total, 4));
        hipLaunchKernelGGL(k_candidate_lists, dim3(T), dim3(BLOCK), 0, stream, p->tiles.ptr, p->masks.ptr, p->tileCandidates.ptr,
                           (const int*) p->listOffsets.ptr, p->lists.ptr);
        IXH_HIP_CHECK(hipGetLastError());
    }
    if (cutoff) {
        p->buildPos.reserve(3*(size_t) n);
        IXH_HIP_CHECK(hipMemcpyAsync(p->buildPos.ptr, p->pos3.ptr, sizeof(double)*3*(size_t) n, hipMemcpyDeviceToDevice, stream));
    }
    // both block frames sized with the list: evaluations between builds allocate nothing
    for (int f = 0; f < 2; f++) {
        p->localFrame[f].reserve(r4*p->padded);
        p->centerFrame[f].reserve(3*(size_t) p->numBlocks);
    }
    IXH_HIP_CHECK(hipStreamSynchronize(stream));
    p->built = true;
}

// Launch a kernel with events bound to its start and completion (hipExtLaunchKernelGGL: the dispatch itself carries
// them, no marker packet enters the queue), or plainly when there are none (also the form a graph capture records).
template <typename Kernel, typename... Args>
void launch_bound(Kernel kernel, dim3 grid, dim3 block, size_t shared, hipStream_t stream, hipEvent_t start, hipEvent_t end,
                  Args... args) {
    if (start != nullptr || end != nullptr)
        hipExtLaunchKernelGGL(kernel, grid, block, (std::uint32_t) shared, stream, start, end, 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, shared, stream, args...);
}

// The tile kernel of a cutoff launch: the fixed table form where the tables allow it (fixed: ixh_pair::fixedForm and
// no index-order dependence), else the general form.
template <typename Real, int T, bool TAB, bool ALCH>
void launch_queue(const TileArgs<Real>& a, bool fixed, int blocks, size_t tableBytes, hipStream_t stream) {
    const dim3 grid(blocks), group(64*QUEUE_GROUP_WAVES);
    if constexpr (TAB) {
        if (fixed) {
            hipLaunchKernelGGL((k_pair_queue<Real, T, TAB, ALCH, true>), grid, group, tableBytes, stream, a);
            return;
        }
    }
    hipLaunchKernelGGL((k_pair_queue<Real, T, TAB, ALCH>), grid, group, tableBytes, stream, a);
}

// Workgroups of a tile launch over numTiles tiles.
int launch_workgroups(const ixh_pair* p, int numTiles) {
    if (p->method != IXH_NO_CUTOFF && p->numItems > 0 && numTiles == p->numTiles)
        return p->numItems/QUEUE_GROUP_WAVES;   // one wave per work item
    const int groupWaves = (p->method != IXH_NO_CUTOFF ? QUEUE_GROUP_WAVES : 4);
    return std::max(1, std::min((numTiles+groupWaves-1)/groupWaves, p->gridLimit));
}

// Energy and pair-count partials of a launch over all of the handle's tiles: one per workgroup (cutoff tiles) or per
// wave (all-pairs tiles).
long long launch_partials(const ixh_pair* p) {
    const long long blocks = launch_workgroups(p, p->numTiles);
    return (p->method != IXH_NO_CUTOFF ? blocks : 4*blocks);
}

// The tile kernel of an evaluation and its reductions.  end: an event to bind to the last kernel (nullptr: none).
// staticScratch: the partial sums go to the static scratch instead of the handle's buffers, and nothing is timed.
// Returns the event that marks the end of the work (end), or nullptr if none was bound.
template <typename Real>
hipEvent_t launch_tiles(ixh_pair* p, const ixh_buffers* b, int terms, int addEnergy, double* segmentEnergies, hipStream_t stream,
                        hipEvent_t end, int subsetTiles, int subsetSegments, void* staticScratch) {
    typedef typename Vec4T<Real>::type R4;
    TileArgs<Real> a;
    a.tiles = p->tiles.ptr;
    a.numTiles = p->numTiles;
    a.partners = p->partners.ptr;
    a.tileSlots = p->tileSlots.ptr;
    a.masks = p->masks.ptr;
    a.listOffsets = p->listOffsets.ptr;
    a.lists = p->lists.ptr;
    a.perm = p->perm.ptr;
    a.local = (const R4*) p->localFrame[p->frame].ptr;
    a.centers = p->centerFrame[p->frame].ptr;
    a.pA = (const R4*) p->sortedA.ptr;
    a.pB = (const R4*) p->sortedB.ptr;
    a.pC = (const R4*) p->sortedC.ptr;
    a.flags = p->sortedFlags.ptr;
    a.n = p->n;
    a.method = p->method;
    a.periodic = (p->method == IXH_CUTOFF_PERIODIC);
    a.box = make_box(b->box);
    a.boxR[0] = Real(b->box[0]);
    a.boxR[1] = Real(b->box[3]);
    a.boxR[2] = Real(b->box[4]);
    a.boxR[3] = Real(b->box[6]);
    a.boxR[4] = Real(b->box[7]);
    a.boxR[5] = Real(b->box[8]);
    a.boxR[6] = Real(a.periodic ? 1.0/b->box[0] : 0.0);
    a.boxR[7] = Real(a.periodic ? 1.0/b->box[4] : 0.0);
    a.boxR[8] = Real(a.periodic ? 1.0/b->box[8] : 0.0);
    a.boxInv[0] = (a.periodic ? 1.0/b->box[0] : 0.0);
    a.boxInv[1] = (a.periodic ? 1.0/b->box[4] : 0.0);
    a.boxInv[2] = (a.periodic ? 1.0/b->box[8] : 0.0);
    a.cutoff = p->cutoff;
    a.cutoff2 = Real(p->cutoff*p->cutoff);
    a.band = Real(sizeof(Real) == 4 ? 1e-5 : 1e-12);
    a.useSwitch = (p->method != IXH_NO_CUTOFF && p->useSwitch);
    a.switchDistance = Real(p->switchDistance);
    a.invSwitchWidth = Real(a.useSwitch ? 1.0/(p->cutoff-p->switchDistance) : 0.0);
    double lambda = (p->anyAlchemical ? b->lambda : 1.0);
    a.lambda = Real(lambda);
    a.useSoftcore = (p->softcoreRadius > 0 && p->anyAlchemical);
    double R3 = p->softcoreRadius*p->softcoreRadius*p->softcoreRadius;
    a.softcoreTerm = Real((1-lambda)*R3*R3);
    a.settings = p->settings;
    a.posq = b->posq;
    a.correction = (const float4*) b->posq_correction;
    a.precision = b->precision;
    a.force = b->force;
    a.stride = b->force_stride;
    const bool tables = (sizeof(Real) == 4 && p->useTables);
    a.cls = p->sortedClasses.ptr;
    a.tables = p->tables.ptr;
    a.tableFloats = p->tableFloats;
    a.numDisp = p->numDisp;
    a.numExch = p->numExch;
    a.numSr = p->numSr;
    a.exchOffset = p->exchOffset;
    a.srOffset = p->srOffset;
    a.highOrder = p->highOrder;
    a.singleBranch = p->singleBranch;
    a.ordered = ((p->terms & IXH_TERM_EXCHANGE) && !p->settings.exchange_scheme.about_smaller) ||
                ((p->terms & IXH_TERM_SHORT_RANGE) && !p->settings.short_range_scheme.about_smaller);
    a.beta2 = (float) (2*p->dampingScale);
    const size_t tableBytes = (tables ? sizeof(float)*p->tableFloats : 0);
    // the fixed form's overlap rows hold the energy of the handle's overlap terms together: a launch takes it when it
    // evaluates all of them (or none)
    const int overlapTerms = IXH_TERM_EXCHANGE | IXH_TERM_SHORT_RANGE;
    const bool fixed = (tables && p->fixedForm && !a.ordered &&
                        ((terms & overlapTerms) == 0 || (terms & overlapTerms) == (p->terms & overlapTerms)));
    a.tileIndex = (p->ordered ? (const int*) p->tileOrder.ptr : (const int*) nullptr);
    a.items = (p->method != IXH_NO_CUTOFF && p->numItems > 0 && subsetTiles < 0 ? (const int2*) p->items.ptr : nullptr);
    if (subsetTiles >= 0) {
        a.numTiles = subsetTiles;
        a.tileIndex = p->subsetTiles.ptr;
    }
    const bool compact = (p->method != IXH_NO_CUTOFF);
    const int blocks = launch_workgroups(p, a.numTiles);
    a.tileEnergy = nullptr;
    if (p->numSegments > 0) {
        p->tileEnergy.reserve(std::max(p->numTiles, 1));
        a.tileEnergy = p->tileEnergy.ptr;
    }
    // partial sums of the launch: one per workgroup (cutoff tiles) or per wave (all-pairs tiles), in the handle's
    // buffers or in the static scratch (whose capacity the caller has checked: launch_partials)
    const int numWaves = (compact ? blocks : 4*blocks);
    if (staticScratch != nullptr) {
        a.waveEnergy = (double*) staticScratch;
        a.waveCount = (int*) ((char*) staticScratch+sizeof(double)*STATIC_PARTIALS);
    }
    else {
        p->waveEnergy.reserve(numWaves);
        p->waveCount.reserve(numWaves);
        a.waveEnergy = p->waveEnergy.ptr;
        a.waveCount = p->waveCount.ptr;
    }
    const bool segments = (segmentEnergies != nullptr && p->numSegments > 0);
    const int segmentCount = (!segments ? 0 : subsetTiles >= 0 ? subsetSegments : p->numSegments);
    hipEvent_t marked = nullptr;
    if (a.numTiles > 0) {
        // with profiling, events recorded around the kernel (events bound to its dispatch would also count the wait
        // for the previous kernel)
        const bool timed = (staticScratch == nullptr && p->timer.start(stream));
        switch (terms) {
#define IXH_LAUNCH(T, TAB, ALCH) \
            if (compact) \
                launch_queue<Real, T, TAB, ALCH>(a, fixed, blocks, tableBytes, stream); \
            else \
                hipLaunchKernelGGL((k_pair_rotate<Real, T, TAB, ALCH>), dim3(blocks), dim3(256), tableBytes, stream, a);
// the variants without alchemical steps are built for the FP32 type-table kernels only
#define IXH_CASE(T) case T: \
            if constexpr (sizeof(Real) == 4) { \
                if (tables) { \
                    if (p->anyAlchemical) { \
                        IXH_LAUNCH(T, true, true) \
                    } \
                    else { \
                        IXH_LAUNCH(T, true, false) \
                    } \
                    break; \
                } \
            } \
            IXH_LAUNCH(T, false, true) \
            break;
            IXH_CASE(1) IXH_CASE(2) IXH_CASE(3) IXH_CASE(4) IXH_CASE(5) IXH_CASE(6) IXH_CASE(7)
#undef IXH_CASE
#undef IXH_LAUNCH
            default:
                throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_execute: illegal terms");
        }
        IXH_HIP_CHECK(hipGetLastError());
        if (timed)
            p->timer.stop(stream);
    }
    launch_bound(k_reduce_waves, dim3(1), dim3(1024), 0, stream, nullptr, segmentCount > 0 ? nullptr : end,
                 (const double*) a.waveEnergy, (const int*) a.waveCount, a.numTiles > 0 ? numWaves : 0, b->energy,
                 p->pairCount.ptr, addEnergy);
    IXH_HIP_CHECK(hipGetLastError());
    marked = end;
    if (segmentCount > 0) {
        launch_bound(k_reduce_segments, dim3(segmentCount), dim3(256), 0, stream, nullptr, end, (const double*) p->tileEnergy.ptr,
                     (const int*) p->segTilesDev.ptr, subsetTiles >= 0 ? (const int*) p->subsetSegments.ptr : (const int*) nullptr,
                     segmentEnergies);
        IXH_HIP_CHECK(hipGetLastError());
        marked = end;
    }
    return marked;
}

// One evaluation (ixh_pair_execute_subset).  staticScratch != nullptr: the no-allocation form
// (ixh_pair_execute_noalloc): the list must exist and still hold, the launch's partial sums go to the static scratch,
// and no device or host memory, event or timer is created.
void execute_pair(ixh_pair* p, const ixh_buffers* b, int terms, int flags, double* segmentEnergies, const int* segments,
                  int numSelected, void* staticScratch) {
    if (numSelected >= 0 && (p == nullptr || p->numSegments == 0 || (numSelected > 0 && segments == nullptr)))
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_execute_subset: a subset needs a segmented handle and a segment list");
    if (p == nullptr || b == nullptr)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_execute: null argument");
    if (b->num_particles != p->n)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_execute: the buffers hold a different number of particles");
    if (b->precision != p->precision && !(b->precision != IXH_DOUBLE && p->precision != IXH_DOUBLE))
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_execute: the buffers have a different precision");
    terms &= p->terms;
    if (terms == 0 || p->n == 0)
        return;
    if (p->method == IXH_CUTOFF_PERIODIC) {
        const double* bx = b->box;
        if (p->cutoff > 0.5*bx[0] || p->cutoff > 0.5*bx[4] || p->cutoff > 0.5*bx[8])
            throw Error(IXH_ERROR_INVALID_ARGUMENT,
                        "InteraxPairForce: The cutoff distance cannot be greater than half the periodic box size.");
    }
    const bool noalloc = (staticScratch != nullptr);
    if (noalloc) {
        if (p->numSegments > 0 || numSelected >= 0)
            throw Error(IXH_ERROR_UNSUPPORTED, "ixh_pair_execute_noalloc: segmented handles are not supported");
        if (!p->built)
            throw Error(IXH_ERROR_STATE, "ixh_pair_execute_noalloc: the handle has no neighbour list yet (its first "
                                         "evaluation must be made by ixh_pair_execute)");
        if (launch_partials(p) > STATIC_PARTIALS)
            throw Error(IXH_ERROR_CAPACITY, "ixh_pair_execute_noalloc: the neighbour list needs " +
                                            std::to_string(launch_partials(p)) + " partial sums; the static scratch holds " +
                                            std::to_string(STATIC_PARTIALS));
    }
    DeviceGuard guard(p->device);
    hipStream_t stream = (hipStream_t) b->stream;
    if (noalloc && p->updatePending && p->updateNeedsHost)
        throw Error(IXH_ERROR_STATE, "ixh_pair_execute_noalloc: a parameter update is not complete (the next ixh_pair_execute "
                                     "completes it)");
    finish_parameters(p, stream);
    const float4* corr = (const float4*) b->posq_correction;
    bool rebuild = !p->built;
    const int periodic = (p->method == IXH_CUTOFF_PERIODIC);
    const size_t r4 = (p->precision == IXH_DOUBLE ? sizeof(double4) : sizeof(float4));
    // The list check reads only this call's positions and the positions of the last build, so it runs before
    // this call waits for the previous one: its answer is back while the previous tiles still run.  The same
    // pass writes this call's block frames into the frame the previous call does not read (the call before it,
    // which used this frame, has finished with it first).
    const bool deferredBefore = p->deferredCheck;
    p->deferredCheck = false;
    const int previousFrame = p->frame;   // frameDone[previousFrame] marks the end of the previous call
    bool framed = false;   // this call's block frames are written
    if (!rebuild && p->method != IXH_NO_CUTOFF) {
        // the previous check, whose flag this one clears, has finished
        if (deferredBefore && stream != p->lastStream)
            IXH_HIP_CHECK(hipStreamWaitEvent(stream, p->flagCopied, 0));
        const int f = p->frame ^ 1;
        if (p->frameDone[f] != nullptr && p->frameStream[f] != stream)
            IXH_HIP_CHECK(hipStreamWaitEvent(stream, p->frameDone[f], 0));
        if (!noalloc) {
            p->localFrame[f].reserve(r4*p->padded);
            p->centerFrame[f].reserve(3*(size_t) p->numBlocks);
        }
        p->frame = f;
        const ListCheck check = {p->buildPos.ptr, 0.25*SKIN*SKIN, p->hostFlag+f, p->hostFlag+(f ^ 1)};
        p->checkedFrame = f;
        // a deferred check carries its completion event on its own dispatch
        const bool defer = (flags & IXH_DEFER_LIST_CHECK) != 0;
        if (defer && p->flagCopied == nullptr)   // created with the handle; kept for handles made before
            IXH_HIP_CHECK(hipEventCreateWithFlags(&p->flagCopied, hipEventDisableTiming));
        hipEvent_t checked = (defer ? p->flagCopied : nullptr);
        if (p->precision == IXH_DOUBLE)
            launch_bound(k_local_positions<double, true>, dim3(grid_for(p->padded, 256)), dim3(256), 0, stream, nullptr, checked,
                         b->posq, corr, b->precision, (const int*) p->perm.ptr, p->padded, periodic, make_box(b->box),
                         p->centerFrame[f].ptr, (double4*) p->localFrame[f].ptr, check, b->atom_index);
        else
            launch_bound(k_local_positions<float, true>, dim3(grid_for(p->padded, 256)), dim3(256), 0, stream, nullptr, checked,
                         b->posq, corr, b->precision, (const int*) p->perm.ptr, p->padded, periodic, make_box(b->box),
                         p->centerFrame[f].ptr, (float4*) p->localFrame[f].ptr, check, b->atom_index);
        IXH_HIP_CHECK(hipGetLastError());
        if (defer)
            p->deferredCheck = true;
        else {
            IXH_HIP_CHECK(hipStreamSynchronize(stream));
            rebuild = (p->hostFlag[f] != 0);
        }
        framed = !rebuild;
    }
    // the handle's buffers are reused by every call: a call on another stream waits for the previous one
    if (p->frameDone[previousFrame] != nullptr && stream != p->lastStream)
        IXH_HIP_CHECK(hipStreamWaitEvent(stream, p->frameDone[previousFrame], 0));
    if (rebuild && noalloc)
        throw Error(IXH_ERROR_STATE, "ixh_pair_execute_noalloc: the neighbour list no longer holds for these positions "
                                     "(ixh_pair_execute rebuilds it)");
    p->lastStream = stream;
    if (rebuild)
        build_list(p, b, stream);
    // without a list check (no cutoff, a new list) the frames are written after the previous call, which may have
    // used either frame, has finished (build_list sizes both frames)
    if (!noalloc) {
        p->localFrame[p->frame].reserve(r4*p->padded);
        p->centerFrame[p->frame].reserve(3*(size_t) p->numBlocks);
    }
    int subsetTileCount = -1;
    if (numSelected >= 0) {
        // the tiles of the selected segments, in segment order
        size_t needed = 0;
        for (int k = 0; k < numSelected; k++) {
            const int g = segments[k];
            if (g < 0 || g >= p->numSegments)
                throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_execute_subset: segment index out of range");
            needed += p->segTiles[g+1]-p->segTiles[g];
        }
        needed += numSelected;
        int slot = 0;
        int* tileList = p->subsetStaging.acquire(needed, slot);
        int count = 0;
        for (int k = 0; k < numSelected; k++)
            for (int t = p->segTiles[segments[k]]; t < p->segTiles[segments[k]+1]; t++)
                tileList[count++] = t;
        std::copy(segments, segments+numSelected, tileList+count);
        p->subsetTiles.reserve(std::max(count, 1));
        p->subsetSegments.reserve(std::max(numSelected, 1));
        if (count > 0)
            IXH_HIP_CHECK(hipMemcpyAsync(p->subsetTiles.ptr, tileList, sizeof(int)*count, hipMemcpyHostToDevice, stream));
        if (numSelected > 0)
            IXH_HIP_CHECK(hipMemcpyAsync(p->subsetSegments.ptr, tileList+count, sizeof(int)*numSelected, hipMemcpyHostToDevice,
                                         stream));
        p->subsetStaging.recorded(slot, stream);
        subsetTileCount = count;
    }
    const ListCheck none = {};
    // the end of the call: an event bound to its last kernel, or recorded after it in a graph capture
    if (p->frameDone[p->frame] == nullptr)
        IXH_HIP_CHECK(hipEventCreateWithFlags(&p->frameDone[p->frame], hipEventDisableTiming));
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    IXH_HIP_CHECK(hipStreamIsCapturing(stream, &capture));
    hipEvent_t end = (capture == hipStreamCaptureStatusNone ? p->frameDone[p->frame] : nullptr);
    hipEvent_t marked = nullptr;
    if (p->precision == IXH_DOUBLE) {
        if (!framed)
            hipLaunchKernelGGL((k_local_positions<double, false>), dim3(grid_for(p->padded, 256)), dim3(256), 0, stream, b->posq,
                               corr, b->precision, p->perm.ptr, p->padded, periodic, make_box(b->box), p->centerFrame[p->frame].ptr,
                               (double4*) p->localFrame[p->frame].ptr, none, b->atom_index);
        IXH_HIP_CHECK(hipGetLastError());
        marked = launch_tiles<double>(p, b, terms, (flags & IXH_INCLUDE_ENERGY) != 0, segmentEnergies, stream, end,
                                      subsetTileCount, std::max(numSelected, 0), staticScratch);
    }
    else {
        if (!framed)
            hipLaunchKernelGGL((k_local_positions<float, false>), dim3(grid_for(p->padded, 256)), dim3(256), 0, stream, b->posq,
                               corr, b->precision, p->perm.ptr, p->padded, periodic, make_box(b->box), p->centerFrame[p->frame].ptr,
                               (float4*) p->localFrame[p->frame].ptr, none, b->atom_index);
        IXH_HIP_CHECK(hipGetLastError());
        marked = launch_tiles<float>(p, b, terms, (flags & IXH_INCLUDE_ENERGY) != 0, segmentEnergies, stream, end,
                                     subsetTileCount, std::max(numSelected, 0), staticScratch);
    }
    if (marked == nullptr)
        IXH_HIP_CHECK(hipEventRecord(p->frameDone[p->frame], stream));
    p->frameStream[p->frame] = stream;
    p->frameUsed[p->frame] = true;
}

}  // namespace

namespace ixh {

void pair_set_grid_limit(ixh_pair* pair, int workgroups) {
    pair->gridLimit = std::max(1, workgroups);
}

int pair_num_tiles(const ixh_pair* pair) {
    return pair->numTiles;
}

}  // namespace ixh

extern "C" {

int ixh_pair_create(int device, const ixh_pair_desc* d, ixh_pair** out) {
    return ixh_pair_create_segmented(device, d, 0, nullptr, out);
}

int ixh_pair_create_segmented(int device, const ixh_pair_desc* d, int num_segments, const int* particle_offsets, ixh_pair** out) {
    return guarded([&] {
        if (out == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_create: null output pointer");
        check_desc(d);
        if (num_segments < 0 || (num_segments > 0 && particle_offsets == nullptr))
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_create_segmented: invalid segments");
        if (num_segments > 0) {
            if (d->method != IXH_NO_CUTOFF)
                throw Error(IXH_ERROR_UNSUPPORTED, "ixh_pair_create_segmented: segments require the NoCutoff method");
            if (particle_offsets[0] != 0 || particle_offsets[num_segments] != d->num_particles)
                throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_create_segmented: the segments must cover the particles");
            for (int g = 0; g < num_segments; g++)
                if (particle_offsets[g+1] < particle_offsets[g])
                    throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_create_segmented: decreasing segment offsets");
        }
        DeviceGuard guard(device);
        ixh_pair* p = new ixh_pair();
        try {
            p->device = device;
            p->n = d->num_particles;
            p->terms = d->terms;
            p->method = d->method;
            p->cutoff = d->cutoff;
            p->useSwitch = (d->method != IXH_NO_CUTOFF && d->use_switching_function);
            p->switchDistance = d->switching_distance;
            p->softcoreRadius = d->softcore_radius;
            p->precision = d->precision;
            p->exchangeMethod = d->exchange_overlap_method;
            p->shortRangeMethod = d->short_range_overlap_method;
            p->useLRC = d->use_long_range_correction;
            p->alchemical.assign(p->n, 0);
            if (d->alchemical != nullptr)
                for (int i = 0; i < p->n; i++)
                    p->alchemical[i] = (d->alchemical[i] != 0);
            p->anyAlchemical = std::any_of(p->alchemical.begin(), p->alchemical.end(), [](unsigned char c) { return c != 0; });
            p->exclusions.assign(d->exclusions, d->exclusions+2*(size_t) d->num_exclusions);
            {
                // per-particle sorted exclusion lists (the sets of ReferenceInteraxPairIxn.cpp:43-49)
                std::vector<std::vector<int>> lists(p->n);
                for (int e = 0; e < d->num_exclusions; e++) {
                    int p1 = d->exclusions[2*e], p2 = d->exclusions[2*e+1];
                    lists[p1].push_back(p2);
                    lists[p2].push_back(p1);
                }
                std::vector<int> offsets(p->n+1, 0), partners;
                for (int i = 0; i < p->n; i++) {
                    std::sort(lists[i].begin(), lists[i].end());
                    lists[i].erase(std::unique(lists[i].begin(), lists[i].end()), lists[i].end());
                    partners.insert(partners.end(), lists[i].begin(), lists[i].end());
                    offsets[i+1] = (int) partners.size();
                }
                p->exclusionOffsets.upload(offsets.data(), offsets.size(), nullptr);
                p->exclusionPartners.upload(partners.data(), std::max<size_t>(partners.size(), 0), nullptr);
                p->exclusionPartners.reserve(std::max<size_t>(partners.size(), 1));
            }
            if (num_segments > 0) {
                p->numSegments = num_segments;
                p->segOffsets.assign(particle_offsets, particle_offsets+num_segments+1);
                p->segBlocks.assign(num_segments+1, 0);
                for (int g = 0; g < num_segments; g++)
                    p->segBlocks[g+1] = p->segBlocks[g]+(particle_offsets[g+1]-particle_offsets[g]+BLOCK-1)/BLOCK;
                p->numBlocks = std::max(1, p->segBlocks[num_segments]);
            }
            else
                p->numBlocks = std::max(1, (p->n+BLOCK-1)/BLOCK);
            p->padded = p->numBlocks*BLOCK;
            hipStream_t stream = nullptr;
            p->pairCount.reserve(1);
            IXH_HIP_CHECK(hipHostMalloc((void**) &p->hostFlag, 2*sizeof(int), hipHostMallocCoherent));
            p->hostFlag[0] = p->hostFlag[1] = 0;
            // the events of every evaluation, and the static scratch of this device, are set up here so that
            // ixh_pair_execute_noalloc creates nothing
            for (hipEvent_t& e : p->frameDone)
                IXH_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            IXH_HIP_CHECK(hipEventCreateWithFlags(&p->flagCopied, hipEventDisableTiming));
            IXH_HIP_CHECK(hipEventCreateWithFlags(&p->updateDone, hipEventDisableTiming));
            IXH_HIP_CHECK(hipHostMalloc((void**) &p->staging, sizeof(ixh_pair::Staging), hipHostMallocDefault));
            (void) static_scratch(device);
            p->mismatch.reserve(1);
            init_parameters(p, d, stream);
        }
        catch (...) {
            if (p->hostFlag != nullptr)
                (void) hipHostFree(p->hostFlag);
            if (p->staging != nullptr)
                (void) hipHostFree(p->staging);
            if (p->updateDone != nullptr)
                (void) hipEventDestroy(p->updateDone);
            delete p;
            throw;
        }
        *out = p;
    });
}

int ixh_pair_destroy(ixh_pair* p) {
    return guarded([&] {
        if (p == nullptr)
            return;
        DeviceGuard guard(p->device);
        (void) hipDeviceSynchronize();
        if (p->hostFlag != nullptr)
            (void) hipHostFree(p->hostFlag);
        if (p->flagCopied != nullptr)
            (void) hipEventDestroy(p->flagCopied);
        if (p->updateDone != nullptr)
            (void) hipEventDestroy(p->updateDone);
        if (p->staging != nullptr)
            (void) hipHostFree(p->staging);
        for (hipEvent_t e : p->frameDone)
            if (e != nullptr)
                (void) hipEventDestroy(e);
        if (p->lambdaHost != nullptr)
            (void) hipHostFree(p->lambdaHost);
        delete p;
    });
}

int ixh_pair_set_parameters(ixh_pair* p, const ixh_pair_desc* d) {
    return guarded([&] {
        if (p == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_set_parameters: null handle");
        check_desc(d);
        // ReferenceInteraxKernels.cpp:98-102, 142-148, 185-191
        if (d->num_particles != p->n)
            throw Error(IXH_ERROR_STATE, "updateParametersInContext: The number of particles has changed");
        check_update(p, d->terms, d->exchange_overlap_method, d->short_range_overlap_method);
        DeviceGuard guard(p->device);
        hipStream_t stream = nullptr;
        order_update(p, stream);
        IXH_HIP_CHECK(hipStreamSynchronize(stream));
        upload_raw(p, d);
        update_parameters(p, d->damping_scale, stream);
        finish_parameters(p, stream);
        finish_long_range(p);
    });
}

int ixh_pair_set_parameters_device(ixh_pair* p, const ixh_pair_device_parameters* d) {
    return guarded([&] {
        if (p == nullptr || d == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_set_parameters_device: null argument");
        if (((p->terms & IXH_TERM_DISPERSION) && d->dispersion == nullptr) ||
            ((p->terms & IXH_TERM_EXCHANGE) && d->exchange == nullptr) ||
            ((p->terms & IXH_TERM_SHORT_RANGE) && d->short_range == nullptr))
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_set_parameters_device: missing parameters of a term of the handle");
        DeviceGuard guard(p->device);
        hipStream_t stream = (hipStream_t) d->stream;
        order_update(p, stream);
        const size_t n = (size_t) p->n;
        if (n > 0) {
            if (p->terms & IXH_TERM_DISPERSION)
                IXH_HIP_CHECK(hipMemcpyAsync(p->rawDisp.ptr, d->dispersion, sizeof(double)*5*n, hipMemcpyDeviceToDevice, stream));
            if (p->terms & IXH_TERM_EXCHANGE)
                IXH_HIP_CHECK(hipMemcpyAsync(p->rawExch.ptr, d->exchange, sizeof(double)*2*n, hipMemcpyDeviceToDevice, stream));
            if (p->terms & IXH_TERM_SHORT_RANGE)
                IXH_HIP_CHECK(hipMemcpyAsync(p->rawSr.ptr, d->short_range, sizeof(double)*3*n, hipMemcpyDeviceToDevice, stream));
        }
        update_parameters(p, (p->terms & IXH_TERM_DISPERSION) ? d->damping_scale : p->dampingScale, stream);
    });
}

int ixh_pair_execute(ixh_pair* p, const ixh_buffers* b, int terms, int flags) {
    return ixh_pair_execute_segments(p, b, terms, flags, nullptr);
}

int ixh_pair_execute_segments(ixh_pair* p, const ixh_buffers* b, int terms, int flags, double* segmentEnergies) {
    return ixh_pair_execute_subset(p, b, terms, flags, segmentEnergies, nullptr, -1);
}

int ixh_pair_execute_subset(ixh_pair* p, const ixh_buffers* b, int terms, int flags, double* segmentEnergies, const int* segments,
                            int numSelected) {
    return guarded([&] { execute_pair(p, b, terms, flags, segmentEnergies, segments, numSelected, nullptr); });
}

int ixh_pair_execute_noalloc(ixh_pair* p, const ixh_buffers* b, int terms, int flags) {
    return guarded([&] {
        if (p == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_execute_noalloc: null argument");
        void* address = nullptr;
        {
            DeviceGuard guard(p->device);
            address = static_scratch(p->device);   // resolved when the handle was created
        }
        execute_pair(p, b, terms, flags, nullptr, nullptr, -1, address);
    });
}

int ixh_pair_long_range_coefficients(ixh_pair* p, double* coefficient, double* alchemical) {
    return guarded([&] {
        if (p == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_long_range_coefficients: null handle");
        {
            DeviceGuard guard(p->device);
            finish_long_range(p);
        }
        if (coefficient != nullptr)
            *coefficient = p->lrcCoefficient;
        if (alchemical != nullptr)
            *alchemical = p->lrcAlchemical;
    });
}

int ixh_pair_long_range_energy(ixh_pair* p, const double* box, double lambda, double* energy) {
    return guarded([&] {
        if (p == nullptr || energy == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_long_range_energy: null argument");
        {
            DeviceGuard guard(p->device);
            finish_long_range(p);
        }
        double e = 0;
        if (p->lrcCoefficient != 0 || p->lrcAlchemical != 0) {
            if (box == nullptr)
                throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_long_range_energy: null box");
            double lam = (p->anyAlchemical ? lambda : 1.0);
            e = (p->lrcCoefficient+lam*p->lrcAlchemical)/(box[0]*box[4]*box[8]);
        }
        *energy = e;
    });
}

}  // extern "C"

namespace {

// Energies at several lambda values, asynchronously on the buffers' stream (the work of ixh_pair_lambda_energies up
// to the copies into the handle's pinned lambdaHost; executeFlags may hold IXH_DEFER_LIST_CHECK).  The caller's
// device must be current, terms must be nonzero and the handle nonempty.
void lambda_launch(ixh_pair* p, const ixh_buffers* b, int terms, const double* lambdas, int count, int executeFlags) {
    hipStream_t stream = (hipStream_t) b->stream;
    // 1. the handle's own evaluation at the buffers' lambda, into scratch forces and energy
    const long long forceWords = 3LL*b->force_stride;
    p->lambdaForce.reserve(forceWords);
    p->lambdaEnergy.reserve(1);
    hipLaunchKernelGGL(k_zero_fixed, dim3(grid_for(forceWords, 256)), dim3(256), 0, stream, p->lambdaForce.ptr, forceWords,
                       p->lambdaEnergy.ptr);
    IXH_HIP_CHECK(hipGetLastError());
    ixh_buffers own = *b;
    own.force = p->lambdaForce.ptr;
    own.energy = p->lambdaEnergy.ptr;
    execute_pair(p, &own, terms, IXH_INCLUDE_FORCES | IXH_INCLUDE_ENERGY | (executeFlags & IXH_DEFER_LIST_CHECK), nullptr, nullptr, -1,
                 nullptr);
    // 2. the pairs of the alchemical particles at every lambda and at the buffers' own
    const int values = count+1;
    if (p->anyAlchemical) {
        LambdaArgs a;
        a.posq = b->posq;
        a.correction = (const float4*) b->posq_correction;
        a.precision = b->precision;
        a.n = p->n;
        a.params = p->lambdaParams.ptr;
        a.flags = p->flags.ptr;
        a.atomIndex = b->atom_index;
        a.alchList = p->alchList.ptr;
        a.exclusionOffsets = p->exclusionOffsets.ptr;
        a.exclusionPartners = p->exclusionPartners.ptr;
        a.method = p->method;
        a.box = make_box(b->box);
        a.cutoff = p->cutoff;
        a.settings = p->settings;
        a.useSwitch = (p->method != IXH_NO_CUTOFF && p->useSwitch);
        a.switchDistance = p->switchDistance;
        a.invSwitchWidth = (a.useSwitch ? 1.0/(p->cutoff-p->switchDistance) : 0.0);
        a.softcoreR6 = std::pow(p->softcoreRadius, 6);
        a.count = values;
        for (int k = 0; k < count; k++)
            a.lambdas[k] = lambdas[k];
        a.lambdas[count] = b->lambda;
        const int chunks = (p->n+LAMBDA_CHUNK-1)/LAMBDA_CHUNK;
        const int rows = p->numAlchemical*chunks;
        p->lambdaOut.reserve((size_t) rows*values+values);
        double* sums = p->lambdaOut.ptr+(size_t) rows*values;
        if (rows > 0) {
            switch (terms) {
#define IXH_LAMBDA_CASE(T) case T: \
                if (p->precision == IXH_DOUBLE) \
                    hipLaunchKernelGGL((k_lambda_energies<double, T>), dim3(rows), dim3(64), 0, stream, a, chunks, p->lambdaOut.ptr); \
                else \
                    hipLaunchKernelGGL((k_lambda_energies<float, T>), dim3(rows), dim3(64), 0, stream, a, chunks, p->lambdaOut.ptr); \
                break;
                IXH_LAMBDA_CASE(1) IXH_LAMBDA_CASE(2) IXH_LAMBDA_CASE(3) IXH_LAMBDA_CASE(4) IXH_LAMBDA_CASE(5) IXH_LAMBDA_CASE(6)
                IXH_LAMBDA_CASE(7)
#undef IXH_LAMBDA_CASE
            }
            IXH_HIP_CHECK(hipGetLastError());
        }
        hipLaunchKernelGGL(k_lambda_columns, dim3(values), dim3(64), 0, stream, p->lambdaOut.ptr, rows, values, sums);
        IXH_HIP_CHECK(hipGetLastError());
    }
    // 3. E(lambda_k) = E(own lambda) - E_alchemical(own lambda) + E_alchemical(lambda_k)
    const size_t words = 1+(p->anyAlchemical ? (size_t) values : 0);
    if (words > p->lambdaHostCount) {
        if (p->lambdaHost != nullptr)
            IXH_HIP_CHECK(hipHostFree(p->lambdaHost));
        p->lambdaHost = nullptr;
        IXH_HIP_CHECK(hipHostMalloc((void**) &p->lambdaHost, sizeof(double)*words, hipHostMallocDefault));
        p->lambdaHostCount = words;
    }
    IXH_HIP_CHECK(hipMemcpyAsync(p->lambdaHost, p->lambdaEnergy.ptr, sizeof(double), hipMemcpyDeviceToHost, stream));
    if (words > 1) {
        const size_t rows = (size_t) p->numAlchemical*((p->n+LAMBDA_CHUNK-1)/LAMBDA_CHUNK);
        IXH_HIP_CHECK(hipMemcpyAsync(p->lambdaHost+1, p->lambdaOut.ptr+rows*values, sizeof(double)*values,
                                     hipMemcpyDeviceToHost, stream));
    }
}

// E(lambda_k) from the copies of lambda_launch, once they have arrived.
void lambda_collect(const ixh_pair* p, int count, double* energies) {
    const bool windows = p->anyAlchemical;
    const double ownEnergy = p->lambdaHost[0];
    for (int k = 0; k < count; k++)
        energies[k] = ownEnergy+(windows ? p->lambdaHost[1+k]-p->lambdaHost[1+count] : 0.0);
}

}  // namespace

namespace ixh {

void pair_lambda_launch(ixh_pair* pair, const ixh_buffers* buffers, const double* lambdas, int count, int flags) {
    if (count < 1 || count > MAX_LAMBDA_WINDOWS)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_lambda_energies: between 1 and 16 lambda values");
    if (pair->n == 0)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_lambda_energies: no particles");
    lambda_launch(pair, buffers, pair->terms, lambdas, count, flags);
}

void pair_lambda_collect(const ixh_pair* pair, int count, double* energies) {
    lambda_collect(pair, count, energies);
}

}  // namespace ixh

extern "C" {

int ixh_pair_lambda_energies(ixh_pair* p, const ixh_buffers* b, int terms, const double* lambdas, int count, double* energies) {
    return guarded([&] {
        if (p == nullptr || b == nullptr || lambdas == nullptr || energies == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_lambda_energies: null argument");
        if (count < 1 || count > MAX_LAMBDA_WINDOWS)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_lambda_energies: between 1 and 16 lambda values");
        if (b->num_particles != p->n)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_lambda_energies: the buffers hold a different number of particles");
        terms &= p->terms;
        std::fill(energies, energies+count, 0.0);
        if (terms == 0 || p->n == 0)
            return;
        DeviceGuard guard(p->device);
        lambda_launch(p, b, terms, lambdas, count, 0);
        IXH_HIP_CHECK(hipStreamSynchronize((hipStream_t) b->stream));
        lambda_collect(p, count, energies);
    });
}

int ixh_pair_kernel_time(ixh_pair* p, double* milliseconds, long long* launches) {
    return guarded([&] {
        if (p == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_kernel_time: null handle");
        DeviceGuard guard(p->device);
        p->timer.collect();
        if (milliseconds != nullptr)
            *milliseconds = p->timer.accumulated;
        if (launches != nullptr)
            *launches = p->timer.launches;
    });
}

int ixh_pair_reset_kernel_time(ixh_pair* p) {
    return guarded([&] {
        if (p == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_reset_kernel_time: null handle");
        DeviceGuard guard(p->device);
        p->timer.reset();
    });
}

int ixh_pair_list_was_valid(ixh_pair* p, int* valid) {
    return guarded([&] {
        if (p == nullptr || valid == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_list_was_valid: null argument");
        *valid = 1;
        if (p->deferredCheck) {
            DeviceGuard guard(p->device);
            IXH_HIP_CHECK(hipEventSynchronize(p->flagCopied));
            *valid = (p->hostFlag[p->checkedFrame] == 0);
        }
    });
}

int ixh_pair_last_pair_count(ixh_pair* p, long long* count) {
    return guarded([&] {
        if (p == nullptr || count == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_last_pair_count: null argument");
        DeviceGuard guard(p->device);
        IXH_HIP_CHECK(hipStreamSynchronize(p->lastStream));
        IXH_HIP_CHECK(hipMemcpy(count, p->pairCount.ptr, sizeof(long long), hipMemcpyDeviceToHost));
    });
}

int ixh_pair_set_tiles_per_item(ixh_pair* p, int tiles_per_item) {
    return guarded([&] {
        if (p == nullptr || tiles_per_item < 1 || tiles_per_item > 64)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_set_tiles_per_item: between 1 and 64 tiles per work item");
        if (p->tilesPerItem != tiles_per_item) {
            DeviceGuard guard(p->device);
            if (p->lastStream != nullptr)
                IXH_HIP_CHECK(hipStreamSynchronize(p->lastStream));
            p->tilesPerItem = tiles_per_item;
            p->built = false;   // the next evaluation lays the list out again
        }
    });
}

int ixh_pair_autotune(ixh_pair* p, const ixh_buffers* b, int terms, int* chosen) {
    return guarded([&] {
        if (p == nullptr || b == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_autotune: null argument");
        if (p->method == IXH_NO_CUTOFF)
            throw Error(IXH_ERROR_UNSUPPORTED, "ixh_pair_autotune: work items are formed for cutoff handles only");
        if (b->num_particles != p->n)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_autotune: the buffers hold a different number of particles");
        DeviceGuard guard(p->device);
        hipStream_t stream = (hipStream_t) b->stream;
        // scratch forces and energy
        const long long forceWords = 3LL*b->force_stride;
        p->lambdaForce.reserve(forceWords);
        p->lambdaEnergy.reserve(1);
        ixh_buffers own = *b;
        own.force = p->lambdaForce.ptr;
        own.energy = p->lambdaEnergy.ptr;
        hipEvent_t start, stop;
        IXH_HIP_CHECK(hipEventCreate(&start));
        IXH_HIP_CHECK(hipEventCreate(&stop));
        int best = p->tilesPerItem;
        float bestMs = 1e30f;
        try {
            for (int shape : {3, 1, 2, 4, 6}) {
                if (p->lastStream != nullptr)
                    IXH_HIP_CHECK(hipStreamSynchronize(p->lastStream));
                p->tilesPerItem = shape;
                p->built = false;
                for (int k = 0; k < 2; k++)   // the list build, then one evaluation with the new list
                    execute_pair(p, &own, terms, IXH_INCLUDE_FORCES | IXH_INCLUDE_ENERGY, nullptr, nullptr, -1, nullptr);
                IXH_HIP_CHECK(hipEventRecord(start, stream));
                for (int k = 0; k < 5; k++)
                    execute_pair(p, &own, terms, IXH_INCLUDE_FORCES | IXH_INCLUDE_ENERGY | IXH_DEFER_LIST_CHECK, nullptr, nullptr,
                                 -1, nullptr);
                IXH_HIP_CHECK(hipEventRecord(stop, stream));
                IXH_HIP_CHECK(hipEventSynchronize(stop));
                float ms = 0;
                IXH_HIP_CHECK(hipEventElapsedTime(&ms, start, stop));
                if (ms < bestMs) {
                    bestMs = ms;
                    best = shape;
                }
            }
        }
        catch (...) {
            (void) hipEventDestroy(start);
            (void) hipEventDestroy(stop);
            throw;
        }
        (void) hipEventDestroy(start);
        (void) hipEventDestroy(stop);
        IXH_HIP_CHECK(hipStreamSynchronize(stream));
        if (best != p->tilesPerItem) {
            p->tilesPerItem = best;
            p->built = false;
        }
        if (chosen != nullptr)
            *chosen = best;
    });
}

int ixh_pair_image_mode(ixh_pair* p, long long* per_block_tiles, long long* tiles) {
    return guarded([&] {
        if (p == nullptr || per_block_tiles == nullptr || tiles == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_image_mode: null argument");
        DeviceGuard guard(p->device);
        if (p->lastStream != nullptr)
            IXH_HIP_CHECK(hipStreamSynchronize(p->lastStream));
        const bool periodic = (p->built && p->method == IXH_CUTOFF_PERIODIC);
        *tiles = p->numTiles;
        *per_block_tiles = (periodic ? p->numTiles-p->pairImageTiles : 0);
    });
}

int ixh_pair_list_statistics(ixh_pair* p, long long* tiles, long long* candidates) {
    return guarded([&] {
        if (p == nullptr || tiles == nullptr || candidates == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_pair_list_statistics: null argument");
        DeviceGuard guard(p->device);
        if (p->lastStream != nullptr)
            IXH_HIP_CHECK(hipStreamSynchronize(p->lastStream));
        *tiles = p->numTiles;
        *candidates = 0;
        if (!p->built || p->numTiles == 0)
            return;
        std::vector<unsigned long long> masks((size_t) p->numTiles*BLOCK);
        std::vector<int2> tileList(p->numTiles);
        IXH_HIP_CHECK(hipMemcpy(masks.data(), p->masks.ptr, sizeof(unsigned long long)*masks.size(), hipMemcpyDeviceToHost));
        IXH_HIP_CHECK(hipMemcpy(tileList.data(), p->tiles.ptr, sizeof(int2)*tileList.size(), hipMemcpyDeviceToHost));
        long long total = 0;
        for (int t = 0; t < p->numTiles; t++) {
            long long c = 0;
            for (int l = 0; l < BLOCK; l++)
                c += __builtin_popcountll(masks[(size_t) t*BLOCK+l]);
            total += (tileList[t].y < 0 ? c/2 : c);
        }
        *candidates = total;
    });
}

}  // extern "C"
