// The InteraxFF Axilrod-Teller-Muto three-body force on gfx950.
//
// Reference behaviour (the contract): ReferenceInteraxThreeBodyIxn.cpp:39-135 (sites, neighbours, triples, force
// sharing), InteraxThreeBodyTerms.h:51-146 (damping and the triple), InteraxThreeBodyForceImpl.h:69-72 (C9) and
// InteraxThreeBodyForceImpl.cpp:42-62 (validation).
//
// Pipeline per evaluation:
//   1. site centroids in FP64, summed in the reference's order, and the neighbour-list check: has a site moved more
//      than half the list skin since the lists were built (interaxThreeBody.cc:11-30 keeps 0.1 nm-skin lists the same
//      way)?
//   2. only when the lists are stale (or new): half neighbour lists within cutoff + skin (sites above, other
//      molecule, decided in FP64 with the reference's minimum image), one wavefront per site with ballot compaction
//      over a cell grid, so each list is in a fixed order;
//   3. one wavefront per site A takes the neighbours of its list within the cutoff at this evaluation's centroids
//      (the reference's FP64 decision), walks their pairs (B, C), tests B-C in FP32 (FP64 near the cutoff or near a
//      change of periodic image) and evaluates the triple in FP32; the forces on B and C go to LDS as 64-bit fixed
//      point (integer atomics, so their order does not matter).  Lists of at most 64 entries take two sites per
//      wavefront (k_triples_packed).  Segmented handles (clusters, k_triple_groups) walk the pairs by blocks of 64
//      neighbours, with both neighbours' forces summed in registers where the blocks are full;
//   4. the site forces are shared equally among the particles of each site.
#include "ixh_common.h"
#include "ixh_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <rocprim/rocprim.hpp>
#include <vector>

using namespace ixh;

namespace {

// Neighbour pairs per work item of the triple kernel of segmented handles (64 steps of 64 pairs for each of its 4
// wavefronts).
constexpr int TRIPLE_CHUNK = 16384;
// Padding of the neighbour lists of plain handles (nm); they are rebuilt when a site has moved skin/2.
constexpr double THREE_BODY_SKIN = 0.1;

struct BoxT {
    double ax, bx, by, cx, cy, cz;
};

// ReferenceInteraxThreeBodyIxn.cpp:74-82, each operation rounded separately as on the host.
__device__ __forceinline__ void ref_delta(const double* from, const double* to, int periodic, const BoxT& b, double d[3]) {
    d[0] = __dsub_rn(to[0], from[0]);
    d[1] = __dsub_rn(to[1], from[1]);
    d[2] = __dsub_rn(to[2], from[2]);
    if (periodic) {
        double s2 = floor(__dadd_rn(__ddiv_rn(d[2], b.cz), 0.5));
        d[0] = __dsub_rn(d[0], __dmul_rn(b.cx, s2));
        d[1] = __dsub_rn(d[1], __dmul_rn(b.cy, s2));
        d[2] = __dsub_rn(d[2], __dmul_rn(b.cz, s2));
        double s1 = floor(__dadd_rn(__ddiv_rn(d[1], b.by), 0.5));
        d[0] = __dsub_rn(d[0], __dmul_rn(b.bx, s1));
        d[1] = __dsub_rn(d[1], __dmul_rn(b.by, s1));
        double s0 = floor(__dadd_rn(__ddiv_rn(d[0], b.ax), 0.5));
        d[0] = __dsub_rn(d[0], __dmul_rn(b.ax, s0));
    }
}

__device__ __forceinline__ double ref_norm2(const double d[3]) {
    return __dadd_rn(__dadd_rn(__dmul_rn(d[0], d[0]), __dmul_rn(d[1], d[1])), __dmul_rn(d[2], d[2]));
}

__device__ __forceinline__ void load_pos(const void* posq, const float4* correction, int precision, int i, double p[3]) {
    if (precision == IXH_DOUBLE) {
        double4 v = ((const double4*) posq)[i];
        p[0] = v.x;
        p[1] = v.y;
        p[2] = v.z;
    }
    else {
        float4 v = ((const float4*) posq)[i];
        p[0] = v.x;
        p[1] = v.y;
        p[2] = v.z;
        if (correction != nullptr) {
            float4 c = correction[i];
            p[0] += c.x;
            p[1] += c.y;
            p[2] += c.z;
        }
    }
}

// 1. ReferenceInteraxThreeBodyIxn.cpp:87-93 (OpenMM's Vec3 divides by multiplying with the reciprocal)
__global__ void k_sites(const void* posq, const float4* correction, int precision, const int* offsets, const int* particles,
                        int numSites, double* sites) {
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= numSites)
        return;
    double s[3] = {0, 0, 0};
    for (int k = offsets[i]; k < offsets[i+1]; k++) {
        double p[3];
        load_pos(posq, correction, precision, particles[k], p);
        s[0] = __dadd_rn(s[0], p[0]);
        s[1] = __dadd_rn(s[1], p[1]);
        s[2] = __dadd_rn(s[2], p[2]);
    }
    double scale = 1.0/(double) (offsets[i+1]-offsets[i]);
    sites[3*i] = __dmul_rn(s[0], scale);
    sites[3*i+1] = __dmul_rn(s[1], scale);
    sites[3*i+2] = __dmul_rn(s[2], scale);
}

// 1'. Plain handles: the centroids (as k_sites), the zeroed site forces, and the neighbour-list check: flag is set
// if a site moved more than sqrt(limit2) since the lists were built (minimum image), and the flag of the other frame is
// cleared for the next call.  Without lists (buildSites == nullptr) there is nothing to check.
__global__ void k_sites_check(const void* posq, const float4* correction, int precision, const int* offsets, const int* particles,
                              int numSites, double* sites, const double* buildSites, int periodic, BoxT box, double limit2, int* flag,
                              int* otherFlag, long long* siteForce) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i == 0)
        *otherFlag = 0;
    if (i < 3*numSites)
        siteForce[i] = 0;
    if (i >= numSites)
        return;
    double s[3] = {0, 0, 0};
    for (int k = offsets[i]; k < offsets[i+1]; k++) {
        double p[3];
        load_pos(posq, correction, precision, particles[k], p);
        s[0] = __dadd_rn(s[0], p[0]);
        s[1] = __dadd_rn(s[1], p[1]);
        s[2] = __dadd_rn(s[2], p[2]);
    }
    const double scale = 1.0/(double) (offsets[i+1]-offsets[i]);
    const double c[3] = {__dmul_rn(s[0], scale), __dmul_rn(s[1], scale), __dmul_rn(s[2], scale)};
    sites[3*i] = c[0];
    sites[3*i+1] = c[1];
    sites[3*i+2] = c[2];
    if (buildSites != nullptr) {
        double d[3];
        ref_delta(buildSites+3*i, c, periodic, box, d);
        if (ref_norm2(d) > limit2)
            *flag = 1;
    }
}

// 2. Half neighbour lists from a cell grid.  Each axis has either one cell (searched alone) or at least three cells
// of at least the cutoff (the 27 neighbouring cells are searched, with periodic wrap or clipped at the grid's edge).
struct CellGrid {
    int n[3];
    int periodic;
    double origin[3];      // nonperiodic: lower corner of the grid
    double scale[3];       // cells per nm along each axis (periodic: cells per box length, applied to x/L)
};

__device__ __forceinline__ int cell_coordinate(double x, int axis, const CellGrid& g) {
    if (g.n[axis] == 1)
        return 0;
    int c;
    if (g.periodic) {
        double f = x*g.scale[axis];      // x / L
        f -= floor(f);
        c = (int) (f*g.n[axis]);
    }
    else
        c = (int) ((x-g.origin[axis])*g.scale[axis]);
    return min(max(c, 0), g.n[axis]-1);
}

__global__ void k_site_cells(const double* sites, int numSites, CellGrid g, unsigned* keys, int* values) {
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= numSites)
        return;
    int cx = cell_coordinate(sites[3*i], 0, g), cy = cell_coordinate(sites[3*i+1], 1, g), cz = cell_coordinate(sites[3*i+2], 2, g);
    keys[i] = (unsigned) ((cz*g.n[1]+cy)*g.n[0]+cx);
    values[i] = i;
}

// Plain handles whose cell grid comes from the box: the centroids, their cells, the zeroed cell ranges and the
// zeroed site forces of an evaluation in one launch (thread i: site i, cell i, force component i).
__global__ void k_prepare_sites(const void* posq, const float4* correction, int precision, const int* offsets, const int* particles,
                                int numSites, double* sites, CellGrid g, unsigned* keys, int* values, int* cellStart, int* cellEnd,
                                int numCells, long long* siteForce, int* cellCount, int* identity) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < numCells) {
        cellStart[i] = 0;
        cellEnd[i] = (identity != nullptr ? numSites : 0);   // one cell: every site, in index order
    }
    if (i < 3*numSites)
        siteForce[i] = 0;
    if (i >= numSites)
        return;
    double s[3] = {0, 0, 0};
    for (int k = offsets[i]; k < offsets[i+1]; k++) {
        double p[3];
        load_pos(posq, correction, precision, particles[k], p);
        s[0] = __dadd_rn(s[0], p[0]);
        s[1] = __dadd_rn(s[1], p[1]);
        s[2] = __dadd_rn(s[2], p[2]);
    }
    const double scale = 1.0/(double) (offsets[i+1]-offsets[i]);
    const double c[3] = {__dmul_rn(s[0], scale), __dmul_rn(s[1], scale), __dmul_rn(s[2], scale)};
    sites[3*i] = c[0];
    sites[3*i+1] = c[1];
    sites[3*i+2] = c[2];
    const int cx = cell_coordinate(c[0], 0, g), cy = cell_coordinate(c[1], 1, g), cz = cell_coordinate(c[2], 2, g);
    const unsigned key = (unsigned) ((cz*g.n[1]+cy)*g.n[0]+cx);
    keys[i] = key;
    values[i] = i;
    if (cellCount != nullptr)
        atomicAdd(&cellCount[key], 1);
    if (identity != nullptr)
        identity[i] = i;
}

// Counting sort of the sites by cell (grids of at most 65 536 cells): start and end of every cell from an exclusive
// scan of the counts (one workgroup, a running total over chunks of 1024), which it leaves zero for the next
// evaluation; then every site takes a place in its cell, and each cell's places are put in site order, so that the
// result is the stable sort by cell.
__global__ void __launch_bounds__(1024) k_cell_scan(int* cellCount, int numCells, int* cellStart, int* cellEnd, int* cursor) {
    // each thread takes 4 consecutive cells of a chunk of 4096
    __shared__ int sum[1024];
    int carry = 0;
    for (int base = 0; base < numCells; base += 4096) {
        const int first = base+4*(int) threadIdx.x;
        int v[4], own = 0;
        for (int k = 0; k < 4; k++) {
            v[k] = (first+k < numCells ? cellCount[first+k] : 0);
            own += v[k];
        }
        sum[threadIdx.x] = own;
        __syncthreads();
        for (int offset = 1; offset < 1024; offset <<= 1) {
            const int add = (threadIdx.x >= (unsigned) offset ? sum[threadIdx.x-offset] : 0);
            __syncthreads();
            sum[threadIdx.x] += add;
            __syncthreads();
        }
        int start = carry+sum[threadIdx.x]-own;
        for (int k = 0; k < 4; k++) {
            const int c = first+k;
            if (c < numCells) {
                cellStart[c] = start;
                cellEnd[c] = start+v[k];
                cursor[c] = start;
                cellCount[c] = 0;
            }
            start += v[k];
        }
        carry += sum[1023];
        __syncthreads();
    }
}

__global__ void k_scatter_sites(const unsigned* keys, int numSites, int* cursor, int* sortedSites) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < numSites)
        sortedSites[atomicAdd(&cursor[keys[i]], 1)] = i;
}

// Each cell's sites in index order: cells of up to SORT_CELL_LDS sites are sorted in a slice of LDS per thread,
// longer ones in place.
constexpr int SORT_CELL_LDS = 32;
__global__ void __launch_bounds__(64) k_sort_cells(const int* cellStart, const int* cellEnd, int numCells, int* sortedSites) {
    __shared__ int slice[64][SORT_CELL_LDS+1];
    const int c = blockIdx.x*blockDim.x+threadIdx.x;
    if (c >= numCells)
        return;
    const int begin = cellStart[c], end = cellEnd[c];
    const int len = end-begin;
    int* v = (len <= SORT_CELL_LDS ? slice[threadIdx.x] : sortedSites+begin);
    if (len <= SORT_CELL_LDS)
        for (int k = 0; k < len; k++)
            v[k] = sortedSites[begin+k];
    for (int k = 1; k < len; k++) {
        const int x = v[k];
        int m = k;
        while (m > 0 && v[m-1] > x) {
            v[m] = v[m-1];
            m--;
        }
        v[m] = x;
    }
    if (len <= SORT_CELL_LDS)
        for (int k = 0; k < len; k++)
            sortedSites[begin+k] = v[k];
}

__global__ void k_iota_int(int* p, int n) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n)
        p[i] = i;
}

__global__ void k_fill_int(int* p, int n, int value) {
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n)
        p[i] = value;
}

// [cellStart, cellEnd) of every cell in the sorted order (empty cells keep 0, 0)
__global__ void k_cell_ranges(const unsigned* sortedKeys, int numSites, int* cellStart, int* cellEnd) {
    int s = blockIdx.x*blockDim.x+threadIdx.x;
    if (s >= numSites)
        return;
    unsigned key = sortedKeys[s];
    if (s == 0 || sortedKeys[s-1] != key)
        cellStart[key] = s;
    if (s == numSites-1 || sortedKeys[s+1] != key)
        cellEnd[key] = s+1;
}

// The reference's minimum image with the divisions replaced by multiplications with the reciprocal box lengths.
// When the cutoff is below half of every box length this gives the same image as ref_delta for every pair within the
// cutoff: the two can differ only where d/L + 1/2 rounds across an integer, i.e. at |d| = L/2 > cutoff.
__device__ __forceinline__ void fast_delta(const double* from, const double* to, int periodic, const BoxT& b, const double inv[3],
                                           double d[3]) {
    d[0] = __dsub_rn(to[0], from[0]);
    d[1] = __dsub_rn(to[1], from[1]);
    d[2] = __dsub_rn(to[2], from[2]);
    if (periodic) {
        double s2 = floor(__dadd_rn(__dmul_rn(d[2], inv[2]), 0.5));
        d[0] = __dsub_rn(d[0], __dmul_rn(b.cx, s2));
        d[1] = __dsub_rn(d[1], __dmul_rn(b.cy, s2));
        d[2] = __dsub_rn(d[2], __dmul_rn(b.cz, s2));
        double s1 = floor(__dadd_rn(__dmul_rn(d[1], inv[1]), 0.5));
        d[0] = __dsub_rn(d[0], __dmul_rn(b.bx, s1));
        d[1] = __dsub_rn(d[1], __dmul_rn(b.by, s1));
        double s0 = floor(__dadd_rn(__dmul_rn(d[0], inv[0]), 0.5));
        d[0] = __dsub_rn(d[0], __dmul_rn(b.ax, s0));
    }
}

// One wave per site i: the sites j > i of another molecule within the cutoff (ReferenceInteraxThreeBodyIxn.cpp:97-103,
// distance with the reference's rounding).  The candidates of the (up to) 27 neighbouring cells are taken as one
// list, 64 at a time.  The hits are written in a fixed order (cells, then site index) to [i capacity, (i+1) capacity);
// count[i] receives their number, so that a list longer than the capacity (its excess not written) is detected and
// the search repeated with more room.
__global__ void k_neighbors(const double* sites, const int* molecule, int numSites, int periodic, BoxT box, double cutoff2, CellGrid g,
                            int exactImage, const int* cellStart, const int* cellEnd, const int* sortedSites, int* count,
                            int capacity, int* nbrIndex) {
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const double* si = sites+3*i;
    const int mi = molecule[i];
    const double inv[3] = {1.0/box.ax, 1.0/box.by, 1.0/box.cz};
    // lane k < 27: the range of neighbouring cell k
    int first = 0, size = 0;
    {
        const int c[3] = {cell_coordinate(si[0], 0, g), cell_coordinate(si[1], 1, g), cell_coordinate(si[2], 2, g)};
        const int span[3] = {g.n[0] == 1 ? 1 : 3, g.n[1] == 1 ? 1 : 3, g.n[2] == 1 ? 1 : 3};
        if (lane < span[0]*span[1]*span[2]) {
            int x = c[0]+(span[0] == 1 ? 0 : lane%3-1);
            int y = c[1]+(span[1] == 1 ? 0 : (lane/span[0])%3-1);
            int z = c[2]+(span[2] == 1 ? 0 : lane/(span[0]*span[1])-1);
            bool inside = true;
            if (g.periodic) {
                x = (x+g.n[0]) % g.n[0];
                y = (y+g.n[1]) % g.n[1];
                z = (z+g.n[2]) % g.n[2];
            }
            else
                inside = (x >= 0 && y >= 0 && z >= 0 && x < g.n[0] && y < g.n[1] && z < g.n[2]);
            if (inside) {
                const int cell = (z*g.n[1]+y)*g.n[0]+x;
                first = cellStart[cell];
                size = cellEnd[cell]-first;
            }
        }
    }
    // inclusive prefix of the range sizes over the lanes
    int prefix = size;
    for (int offset = 1; offset < 64; offset <<= 1) {
        int v = __shfl_up(prefix, offset, 64);
        if (lane >= offset)
            prefix += v;
    }
    const int total = __shfl(prefix, 63, 64);
    int written = 0;
    for (int c0 = 0; c0 < total; c0 += 64) {
        const int c = c0+lane;
        int j = -1;
        // the cell of candidate c: the first lane whose inclusive prefix exceeds c
        int k = 0;
        for (int step = 32; step > 0; step >>= 1) {
            int probe = __shfl(prefix, k+step-1, 64);
            if (probe <= c)
                k += step;
        }
        const int kFirst = __shfl(first, k, 64), kPrefix = __shfl(prefix, k, 64), kSize = __shfl(size, k, 64);
        if (c < total)
            j = sortedSites[kFirst+(c-(kPrefix-kSize))];
        bool hit = false;
        double d[3] = {0, 0, 0};
        if (j > i && molecule[j] != mi) {
            if (exactImage)
                ref_delta(si, sites+3*j, periodic, box, d);
            else
                fast_delta(si, sites+3*j, periodic, box, inv, d);
            hit = ref_norm2(d) < cutoff2;
        }
        const unsigned long long ballot = __ballot(hit);
        const int pos = written+__popcll(ballot & ((1ULL << lane)-1));
        if (hit && pos < capacity)
            nbrIndex[(size_t) i*capacity+pos] = j;
        written += __popcll(ballot);
    }
    if (lane == 0)
        count[i] = written;
}

__global__ void k_strided_starts(int* start, int n, int capacity) {
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n)
        start[i] = i*capacity;
}

template <typename Real>
struct Vec4Of;
template <>
struct Vec4Of<float> {
    typedef float4 type;
};
template <>
struct Vec4Of<double> {
    typedef double4 type;
};

template <typename Real>
struct TripleArgs {
    const double* sites;
    const int* molecule;
    const double* siteParams;    // alpha, frequency, width, alchemical per site
    const int* start;
    const int* count;            // segmented handles: neighbours of site i = [start[i], start[i]+count[i])
    const int* itemSite;         // work items: the site of item w (nullptr: item w is site w, all of its pairs)
    const int* itemFirst;        // work items: the first neighbour pair of item w
    const int2* itemShares;      // work items: the number of shares of item w and their size (pairs)
    const int* itemShareBase;    // work items: the index of item w's first share among all shares
    const int* itemIndex;        // subsets: the item of every workgroup (nullptr: item = workgroup index)
    const int* nbrIndex;
    int numSites;
    int periodic;
    BoxT box;
    Real boxR[9];                // ax, bx, by, cx, cy, cz, 1/ax, 1/by, 1/cz
    Real cutoff2r;
    double cutoff2;
    Real band;                   // relative band around the cutoff and the image boundaries decided in FP64
    Real scaleFactor;
    Real lambda;
    int useDamping;
    Real dampingScale;
    Real invDampingRadius;       // 0: no centroid damping
    int capacity;                // LDS entries
    int skinLists;               // plain handles: the lists hold the sites within cutoff + skin, the neighbours within the
                                 // cutoff are found at each evaluation
    int exactImage;              // skin lists: the reference's minimum image with divisions (small boxes)
    int bcSameImage;             // no periodic boundaries, or every box length above 4 cutoffs: BC = AC - AB needs no
                                 // image (|AC - AB| < 2 cutoffs < half a box length)
    long long* contribution;     // skin lists: the force on the site of list entry e from the triples of the list's
                                 // owner, component k at [3 e + k] (summed by k_gather_site_forces)
    long long* siteForce;        // 3 x numSites fixed point
    double* itemEnergy;          // energy and triple count of every share (plain handles: of every site)
    int* itemCount;
};

// InteraxThreeBodyTerms.h:51-65
template <typename Real>
__device__ __forceinline__ void tt3(Real x, Real& f, Real& df) {
    Real expX = fast_exp2(-x*Real(1.4426950408889634));
    Real term3 = x*x*x*Real(1.0/6.0);
    df = expX*term3;
    if (x < Real(2)) {
        Real term = term3, tail = 0;
        for (int k = 4; k <= 30; k++) {
            term *= x*Real(1.0/k);
            tail += term;
        }
        f = expX*tail;
    }
    else
        f = Real(1)-expX*(Real(1)+x+x*x*Real(0.5)+term3);
}

// 3. Plain handles with lists longer than 64 entries: one wavefront per site A and all of its neighbour pairs
// (shorter lists: k_triples_packed; segmented handles: k_triple_groups).
template <typename Real>
__global__ void __launch_bounds__(64) k_triples(TripleArgs<Real> t) {
    typedef typename Vec4Of<Real>::type R4;
    extern __shared__ double sharedStorage[];
    R4* sDelta = (R4*) sharedStorage;           // AB vector of each neighbour B and its square |AB|^2
    R4* sParam = sDelta+t.capacity;             // alpha w, w, w_A + w and +-1/|AB|^2 (w: the frequency; negative:
                                                // B is alchemical)
    unsigned long long* sForce = (unsigned long long*) (sParam+t.capacity);   // 3 per neighbour
    int* sIndex = (int*) (sForce+3*(size_t) t.capacity);
    int* sMol = sIndex+t.capacity;              // molecule
    Real* sWidth = (Real*) (sMol+t.capacity);   // damping width
    int* sSlot = (int*) (sWidth+t.capacity);    // skin lists: the LDS slot of each list entry, -1 if beyond the cutoff
    const int i = (t.itemIndex != nullptr ? t.itemIndex[blockIdx.x] : (int) blockIdx.x);
    const int lane = threadIdx.x & 63;
    const int begin = t.start[i], listLength = (t.count != nullptr ? t.count[i] : t.start[i+1]-begin);
    const Real aI = Real(t.siteParams[4*i]), wI = Real(t.siteParams[4*i+1]), sI = Real(t.siteParams[4*i+2]);
    const bool alchI = t.siteParams[4*i+3] != 0;
    // the per-neighbour factors of the triple's C9 and geometry, computed once per work item
    auto stage = [&](int a, int j, const double d[3]) {
        const Real x = Real(d[0]), y = Real(d[1]), z = Real(d[2]);
        const Real u = x*x+y*y+z*z;
        R4 dr = {x, y, z, u};
        sDelta[a] = dr;
        const Real alpha = Real(t.siteParams[4*j]), w = Real(t.siteParams[4*j+1]);
        R4 pr = {alpha*w, w, wI+w, (t.siteParams[4*j+3] != 0 ? -fast_rcp(u) : fast_rcp(u))};
        sParam[a] = pr;
        sIndex[a] = j;
        sMol[a] = t.molecule[j];
        sWidth[a] = Real(t.siteParams[4*j+2]);
        sForce[3*a] = sForce[3*a+1] = sForce[3*a+2] = 0;
    };
    int m = listLength;
    {
        // the list's sites within the cutoff now, in list order (ReferenceInteraxThreeBodyIxn.cpp:97-103, the
        // reference's FP64 rounding), compacted by ballot
        const double* si = t.sites+3*i;
        const double inv[3] = {1.0/t.box.ax, 1.0/t.box.by, 1.0/t.box.cz};
        int kept = 0;
        for (int a0 = 0; a0 < listLength; a0 += 64) {
            const int a = a0+lane;
            int j = -1;
            double d[3] = {0, 0, 0};
            bool hit = false;
            if (a < listLength) {
                j = t.nbrIndex[begin+a];
                if (t.exactImage)
                    ref_delta(si, t.sites+3*j, t.periodic, t.box, d);
                else
                    fast_delta(si, t.sites+3*j, t.periodic, t.box, inv, d);
                hit = ref_norm2(d) < t.cutoff2;
            }
            const unsigned long long ballot = __ballot(hit);
            const int slot = kept+__popcll(ballot & ((1ULL << lane)-1));
            if (hit)
                stage(slot, j, d);
            if (a < listLength)
                sSlot[a] = (hit ? slot : -1);
            kept += __popcll(ballot);
        }
        m = kept;
    }
    __syncthreads();
    const Real cI = t.scaleFactor*Real(1.5)*aI*wI;
    Real fA[3] = {0, 0, 0};
    double energy = 0;
    int count = 0;
    // pair indices fit 32 bits: the LDS holds at most a few thousand neighbours
    const int pEnd = m*(m-1)/2;
    // Pair p of the site's neighbour pairs.  The first T = M(M-1)/2 (M = min(m, 64)), which are all the pairs within
    // the first 64 neighbours, are taken in rotation order: rotation s holds the pairs (x, (x+s+1) mod M), x < M (the
    // last, for even M, only x < M/2), so that a step's lanes hold each neighbour about twice (exactly twice, once on
    // each side, when M = 64) instead of up to 63 lanes on one LDS force address.  The energy and forces of a triple
    // do not depend on the order of its two neighbours.  The rest, rows b >= 64, in row order p = b(b-1)/2 + a,
    // where a step's lanes share one b or two consecutive ones.
    const int M = min(m, 64);
    const int T = M*(M-1)/2;
    // The wavefront takes the pairs 64 per step, so that in the rows its steps follow each other along a row: the
    // pair 64 further on is in the same row or the next (rows have b >= 64 pairs), and the forces on the row's b are
    // summed per lane over the steps of the row and over the wavefront only when the wavefront leaves it.
    int R = -1;                 // the row whose forces accR holds (accR1: row R + 1); wavefront-uniform
    bool usedR1 = false;        // some lane has added to accR1
    Real accR[3] = {0, 0, 0}, accR1[3] = {0, 0, 0};
    auto flushRow = [&](int row, Real (&acc)[3]) {
        for (int k = 0; k < 3; k++) {
            const Real sum = wave_sum(acc[k]);
            if (lane == 0)
                atomicAdd(&sForce[3*row+k], (unsigned long long) to_fixed(sum));
            acc[k] = 0;
        }
    };
    {
        const int myBegin = 0, myEnd = pEnd;
        int a = 0, b = 0;
        for (int p0 = myBegin; p0 < myEnd; p0 += 64) {
            const int p = p0+lane;
            if (p0 != myBegin && p0-64 >= T) {
                // the previous step was all rows (b >= 64): advance by 64 pairs, at most into the next row
                a += 64;
                if (a >= b) {
                    a -= b;
                    b++;
                }
            }
            else if (p < T) {
                const int q = (int) p;
                int r = (int) ((float) q*(1.0f/(float) M));
                while (r > 0 && r*M > q)
                    r--;
                while ((r+1)*M <= q)
                    r++;
                a = q-r*M;
                b = a+r+1;
                if (b >= M)
                    b -= M;
            }
            else {
                b = (int) ((1.0f+__builtin_amdgcn_sqrtf(1.0f+8.0f*(float) p))*0.5f);
                while (b*(b-1)/2 > p)
                    b--;
                while ((b+1)*b/2 <= p)
                    b++;
                a = p-b*(b-1)/2;
            }
            const bool active = (p < myEnd);
            // Rows of 64 pairs or more: a step's active lanes (lane 0 always is one) share b or b and b + 1, whose forces
            // go to accR and accR1 instead of up to 64 atomics on one LDS address; the step that straddles the two
            // orders, and the rotation order, add per lane.
            const int bFirst = __builtin_amdgcn_readfirstlane(b);
            const bool rows = (p0 >= T) && __all(!active || b == bFirst || b == bFirst+1);
            if (rows && bFirst != R) {
                if (R >= 0) {
                    flushRow(R, accR);
                    if (bFirst == R+1) {
                        for (int k = 0; k < 3; k++) {
                            accR[k] = accR1[k];
                            accR1[k] = 0;
                        }
                    }
                    else if (usedR1)
                        flushRow(R+1, accR1);
                }
                R = bFirst;
                usedR1 = false;
            }
            if (rows)
                usedR1 = usedR1 || __any(active && b != R);
            if (active) {
                bool ok = (sMol[a] != sMol[b]);
                R4 va = sDelta[a], vb = sDelta[b];
                Real c[3] = {vb.x-va.x, vb.y-va.y, vb.z-va.z};
                if (ok) {
                    bool exact = false;
                    if (t.periodic) {
                        // the reference's minimum image of BC; near a change of image the choice is made in FP64
                        Real u2 = c[2]*t.boxR[8]+Real(0.5);
                        Real s2 = floor(u2);
                        c[0] -= s2*t.boxR[3];
                        c[1] -= s2*t.boxR[4];
                        c[2] -= s2*t.boxR[5];
                        Real u1 = c[1]*t.boxR[7]+Real(0.5);
                        Real s1 = floor(u1);
                        c[0] -= s1*t.boxR[1];
                        c[1] -= s1*t.boxR[2];
                        Real u0 = c[0]*t.boxR[6]+Real(0.5);
                        Real s0 = floor(u0);
                        c[0] -= s0*t.boxR[0];
                        Real f2 = u2-s2, f1 = u1-s1, f0 = u0-s0;
                        exact = (f2 < t.band || f2 > 1-t.band || f1 < t.band || f1 > 1-t.band || f0 < t.band || f0 > 1-t.band);
                    }
                    Real r2 = c[0]*c[0]+c[1]*c[1]+c[2]*c[2];
                    if (exact || fabs(r2-t.cutoff2r) < t.band*t.cutoff2r) {
                        double d[3];
                        ref_delta(t.sites+3*sIndex[a], t.sites+3*sIndex[b], t.periodic, t.box, d);
                        ok = ref_norm2(d) < t.cutoff2;
                        c[0] = Real(d[0]);
                        c[1] = Real(d[1]);
                        c[2] = Real(d[2]);
                    }
                    else
                        ok = r2 < t.cutoff2r;
                }
                R4 pa = sParam[a], pb = sParam[b];
                Real c9 = 0;
                if (ok) {
                    // ReferenceInteraxThreeBodyIxn.cpp:116-120 and InteraxThreeBodyForceImpl.h:69-72:
                    // C9 = 1.5 a_A a_B a_C w_A w_B w_C (w_A+w_B+w_C) / ((w_A+w_B)(w_B+w_C)(w_C+w_A))
                    c9 = cI*pa.x*pb.x*(pa.z+pb.y)*fast_rcp(pa.z*(pa.y+pb.y)*pb.z);
                    if (alchI || pa.w < 0 || pb.w < 0)
                        c9 *= t.lambda;
                    ok = (c9 != 0);
                }
                if (ok) {
                    // InteraxThreeBodyTerms.h:97-146 with a = AB, b = AC, c = BC
                    const Real A[3] = {va.x, va.y, va.z}, B[3] = {vb.x, vb.y, vb.z};
                    const Real u = va.w, v = vb.w;
                    Real w = c[0]*c[0]+c[1]*c[1]+c[2]*c[2];
                    Real pp = A[0]*B[0]+A[1]*B[1]+A[2]*B[2];
                    Real q = A[0]*c[0]+A[1]*c[1]+A[2]*c[2];
                    Real s = B[0]*c[0]+B[1]*c[1]+B[2]*c[2];
                    // FP32: v_rcp_f32 and v_rsq_f32 (an ulp or two) in place of correctly rounded division and sqrt
                    // one reciprocal square root for (uvw)^-1/2, (uvw)^-1, (uvw)^-3/2 and 1/w = uv/(uvw)
                    const Real uvw = u*v*w;
                    const Real rs = fast_rsqrt(uvw);
                    Real invP = rs*rs;
                    Real invP32 = invP*rs;
                    Real invP52 = invP32*invP;
                    Real pqs = pp*q*s;
                    Real e = c9*(invP32-3*pqs*invP52);
                    Real radial = c9*(-3*invP32+15*pqs*invP52);
                    Real cross = -3*c9*invP52;
                    Real ru = radial*fabs(pa.w), rv = radial*fabs(pb.w), rw = radial*(invP*(u*v));
                    // With X = pp s cross, Y = pp q cross, Z = q s cross the gradients with respect to A, B and C = B - A are
                    // dA = A ru + B Z + C X, dB = B rv + A Z + C Y, dC = C rw + A X + B Y; the forces on the neighbours
                    // a (dC - dA) and b (-(dB + dC)) and the sum dA + dB for A are combinations of A, B, C with these
                    // coefficients.
                    const Real X = pp*s*cross, Y = pp*q*cross, Z = q*s*cross;
                    Real aA = X-ru, aB = Y-Z, aC = rw-X;
                    Real bA = -(Z+X), bB = -(rv+Y), bC = -(Y+rw);
                    Real sA = ru+Z, sB = Z+rv, sC = X+Y;
                    if (t.useDamping || t.invDampingRadius > 0) {
                        Real len[3] = {sqrt(u), sqrt(v), sqrt(w)};
                        const Real widthA = sWidth[a], widthB = sWidth[b];
                        Real rates[3] = {2*t.dampingScale/(sI+widthA), 2*t.dampingScale/(sI+widthB), 2*t.dampingScale/(widthA+widthB)};
                        Real f[3], dfdr[3];
                        for (int k = 0; k < 3; k++) {
                            if (t.useDamping) {
                                tt3<Real>(rates[k]*len[k], f[k], dfdr[k]);
                                dfdr[k] *= rates[k];
                            }
                            else {
                                f[k] = 1;
                                dfdr[k] = 0;
                            }
                            if (t.invDampingRadius > 0) {
                                // InteraxThreeBodyTerms.h:71-77
                                Real x = len[k]*t.invDampingRadius;
                                Real x6 = x*x*x;
                                x6 *= x6;
                                Real g = -expm1(-x6);
                                Real dg = 6*x6*exp(-x6)/len[k];
                                dfdr[k] = dfdr[k]*g + f[k]*dg;
                                f[k] *= g;
                            }
                        }
                        // damped gradients: dA' = dA f + A sa, dB' = dB f + B sb, dC' = dC f + C sc
                        Real damping = f[0]*f[1]*f[2];
                        Real sa = e*dfdr[0]*f[1]*f[2]/len[0];
                        Real sb = e*f[0]*dfdr[1]*f[2]/len[1];
                        Real sc = e*f[0]*f[1]*dfdr[2]/len[2];
                        aA = aA*damping-sa;
                        aB *= damping;
                        aC = aC*damping+sc;
                        bA *= damping;
                        bB = bB*damping-sb;
                        bC = bC*damping-sc;
                        sA = sA*damping+sa;
                        sB = sB*damping+sb;
                        sC *= damping;
                        e *= damping;
                    }
                    energy += (double) e;
                    count++;
                    for (int k = 0; k < 3; k++) {
                        fA[k] += A[k]*sA+B[k]*sB+c[k]*sC;
                        atomicAdd(&sForce[3*a+k], (unsigned long long) to_fixed(A[k]*aA+B[k]*aB+c[k]*aC));
                        const Real forceB = A[k]*bA+B[k]*bB+c[k]*bC;
                        if (!rows)
                            atomicAdd(&sForce[3*b+k], (unsigned long long) to_fixed(forceB));
                        else if (b == R)
                            accR[k] += forceB;
                        else
                            accR1[k] += forceB;
                    }
                }
            }
        }
        if (R >= 0) {
            flushRow(R, accR);
            if (usedR1)
                flushRow(R+1, accR1);
        }
    }
    __syncthreads();
    // the neighbour forces as plain stores, one per list entry (zero beyond the cutoff), summed per site by
    // k_gather_site_forces; the force on A is this workgroup's alone (integer sums: exact in any order)
    for (int a = lane; a < listLength; a += 64) {
        const int slot = sSlot[a];
        for (int k = 0; k < 3; k++)
            t.contribution[3*((long long) begin+a)+k] = (slot >= 0 ? (long long) sForce[3*slot+k] : 0LL);
    }
    for (int k = 0; k < 3; k++) {
        long long v = to_fixed(fA[k]);
        for (int offset = 32; offset > 0; offset >>= 1)
            v += __shfl_xor(v, offset, 64);
        if (lane == 0)
            t.siteForce[(size_t) k*t.numSites+i] = v;
    }
    // energy and triple count of the site
    const double e = wave_sum(energy);
    const int c = (int) wave_sum((double) count);
    if (lane == 0) {
        t.itemEnergy[i] = e;
        t.itemCount[i] = c;
    }
}

// The energy and the forces on a and b of one triple (A, a, b) of a plain handle (neighbours staged as k_triples
// stages them): BC with the reference's minimum image, decided in FP64 near the cutoff or a change of image; the same
// arithmetic as k_triples.  Returns false for a triple with no contribution.
template <typename Real, bool DAMPED>
__device__ __forceinline__ bool plain_triple(const TripleArgs<Real>& t, const typename Vec4Of<Real>::type& va,
                                             const typename Vec4Of<Real>::type& pa, Real widthA, int siteA,
                                             const typename Vec4Of<Real>::type& vb, const typename Vec4Of<Real>::type& pb,
                                             Real widthB, int siteB, Real cI, Real sI, bool alchI, Real& energy, Real (&fa)[3],
                                             Real (&fb)[3]) {
    Real c[3] = {vb.x-va.x, vb.y-va.y, vb.z-va.z};
    bool exact = false;
    if (!t.bcSameImage) {
        Real u2 = c[2]*t.boxR[8]+Real(0.5);
        Real s2 = floor(u2);
        c[0] -= s2*t.boxR[3];
        c[1] -= s2*t.boxR[4];
        c[2] -= s2*t.boxR[5];
        Real u1 = c[1]*t.boxR[7]+Real(0.5);
        Real s1 = floor(u1);
        c[0] -= s1*t.boxR[1];
        c[1] -= s1*t.boxR[2];
        Real u0 = c[0]*t.boxR[6]+Real(0.5);
        Real s0 = floor(u0);
        c[0] -= s0*t.boxR[0];
        Real f2 = u2-s2, f1 = u1-s1, f0 = u0-s0;
        exact = (f2 < t.band || f2 > 1-t.band || f1 < t.band || f1 > 1-t.band || f0 < t.band || f0 > 1-t.band);
    }
    Real w = c[0]*c[0]+c[1]*c[1]+c[2]*c[2];
    bool ok;
    if (exact || fabs(w-t.cutoff2r) < t.band*t.cutoff2r) {
        double d[3];
        ref_delta(t.sites+3*siteA, t.sites+3*siteB, t.periodic, t.box, d);
        ok = ref_norm2(d) < t.cutoff2;
        c[0] = Real(d[0]);
        c[1] = Real(d[1]);
        c[2] = Real(d[2]);
        w = c[0]*c[0]+c[1]*c[1]+c[2]*c[2];
    }
    else
        ok = w < t.cutoff2r;
    if (!ok)
        return false;
    // staged parameters (alpha_j w_j / (w_A + w_j), w_j, w_A + w_j, +-1/|Aj|^2), as k_triple_groups stages them
    Real c9 = cI*pa.x*pb.x*(pa.z+pb.y)*fast_rcp(pa.y+pb.y);
    if (t.lambda != Real(1) && (alchI || pa.w < 0 || pb.w < 0))
        c9 *= t.lambda;
    if (c9 == 0)
        return false;
    const Real A[3] = {va.x, va.y, va.z}, B[3] = {vb.x, vb.y, vb.z};
    const Real u = va.w, v = vb.w;
    const Real pp = A[0]*B[0]+A[1]*B[1]+A[2]*B[2];
    const Real q = A[0]*c[0]+A[1]*c[1]+A[2]*c[2];
    const Real s = B[0]*c[0]+B[1]*c[1]+B[2]*c[2];
    const Real uv = u*v;
    const Real rs = fast_rsqrt(uv*w);
    const Real invP = rs*rs;
    const Real invP32 = invP*rs;
    const Real invP52 = invP32*invP;
    const Real pqs = pp*q*s;
    Real e = c9*(invP32-3*pqs*invP52);
    const Real radial = c9*(-3*invP32+15*pqs*invP52);
    const Real cross = -3*c9*invP52;
    const Real ru = radial*fabs(pa.w), rv = radial*fabs(pb.w), rw = radial*(invP*uv);
    const Real X = pp*s*cross, Y = pp*q*cross, Z = q*s*cross;
    Real aA = X-ru, aB = Y-Z, aC = rw-X;
    Real bA = -(Z+X), bB = -(rv+Y), bC = -(Y+rw);
    if constexpr (DAMPED) {
        Real len[3] = {sqrt(u), sqrt(v), sqrt(w)};
        Real rates[3] = {2*t.dampingScale/(sI+widthA), 2*t.dampingScale/(sI+widthB), 2*t.dampingScale/(widthA+widthB)};
        Real f[3], dfdr[3];
        for (int k = 0; k < 3; k++) {
            if (t.useDamping) {
                tt3<Real>(rates[k]*len[k], f[k], dfdr[k]);
                dfdr[k] *= rates[k];
            }
            else {
                f[k] = 1;
                dfdr[k] = 0;
            }
            if (t.invDampingRadius > 0) {
                Real x = len[k]*t.invDampingRadius;
                Real x6 = x*x*x;
                x6 *= x6;
                Real g = -expm1(-x6);
                Real dg = 6*x6*exp(-x6)/len[k];
                dfdr[k] = dfdr[k]*g + f[k]*dg;
                f[k] *= g;
            }
        }
        const Real damping = f[0]*f[1]*f[2];
        const Real sa = e*dfdr[0]*f[1]*f[2]/len[0];
        const Real sb = e*f[0]*dfdr[1]*f[2]/len[1];
        const Real sc = e*f[0]*f[1]*dfdr[2]/len[2];
        aA = aA*damping-sa;
        aB *= damping;
        aC = aC*damping+sc;
        bA *= damping;
        bB = bB*damping-sb;
        bC = bC*damping-sc;
        e *= damping;
    }
    energy = e;
    for (int k = 0; k < 3; k++) {
        fa[k] = A[k]*aA+B[k]*aB+c[k]*aC;
        fb[k] = A[k]*bA+B[k]*bB+c[k]*bC;
    }
    return true;
}

// Sum over the GROUP lanes of each lane group (GROUP a power of two up to 64), every lane of a group receiving the
// same order-fixed total.
template <int GROUP, typename T>
__device__ __forceinline__ T group_sum(T v) {
    for (int offset = GROUP/2; offset > 0; offset >>= 1)
        v += __shfl_xor(v, offset, 64);
    return v;
}

constexpr int PACKED_GROUP = 32;

// 3a. Plain handles whose lists hold at most 64 entries: 64 / GROUP sites per wavefront, GROUP lanes each, so that
// the fixed costs of a site (filtering its list in FP64, staging, the contributions and reductions) are shared out
// over several sites (a site's list holds a few tens of entries, its ~100 pairs a fraction of a wavefront's work).
// Each site has its own LDS region; its pairs are taken in rotation order (rotation r: (x, (x+r+1) mod m)), GROUP
// per step, the forces on a and b added to LDS in fixed point; the force on A is the negated sum of its
// neighbours' (exact), formed by k_gather_site_forces from the per-entry contributions.
template <typename Real, int GROUP, bool DAMPED>
__global__ void __launch_bounds__(64) k_triples_packed(TripleArgs<Real> t, int numItems) {
    typedef typename Vec4Of<Real>::type R4;
    constexpr int SITES = 64/GROUP;
    extern __shared__ double sharedStorage[];
    const int lane = threadIdx.x, g = lane/GROUP, l = lane%GROUP;
    const int cap = t.capacity;
    R4* sDelta = (R4*) sharedStorage+(size_t) g*cap;
    R4* sParam = (R4*) sharedStorage+(size_t) (SITES+g)*cap;
    unsigned long long* sForce = (unsigned long long*) ((R4*) sharedStorage+(size_t) 2*SITES*cap)+(size_t) 3*g*cap;
    int* sIndexAll = (int*) ((unsigned long long*) ((R4*) sharedStorage+(size_t) 2*SITES*cap)+(size_t) 3*SITES*cap);
    int* sIndex = sIndexAll+(size_t) g*cap;
    int* sMol = sIndexAll+(size_t) (SITES+g)*cap;
    int* sSlot = sIndexAll+(size_t) (2*SITES+g)*cap;
    Real* sWidth = (Real*) (sIndexAll+(size_t) 4*SITES*cap)+(size_t) g*cap;
    const int item = SITES*(int) blockIdx.x+g;
    const bool valid = (item < numItems);
    const int i = (valid ? (t.itemIndex != nullptr ? t.itemIndex[item] : item) : 0);
    const int begin = (valid ? t.start[i] : 0);
    const int listLength = (valid ? (t.count != nullptr ? t.count[i] : t.start[i+1]-begin) : 0);
    const Real aI = Real(t.siteParams[4*i]), wI = Real(t.siteParams[4*i+1]), sI = Real(t.siteParams[4*i+2]);
    const bool alchI = t.siteParams[4*i+3] != 0;
    // the list's sites within the cutoff now, in list order (ReferenceInteraxThreeBodyIxn.cpp:97-103, the reference's
    // FP64 rounding), compacted by ballot within the group
    const double* si = t.sites+3*i;
    const double inv[3] = {1.0/t.box.ax, 1.0/t.box.by, 1.0/t.box.cz};
    const unsigned long long groupMask = (GROUP == 64 ? ~0ULL : ((1ULL << GROUP)-1));
    int m = 0;
    for (int a0 = 0; __any(a0 < listLength); a0 += GROUP) {
        const int a = a0+l;
        int j = -1;
        double d[3] = {0, 0, 0};
        bool hit = false;
        if (a < listLength) {
            j = t.nbrIndex[begin+a];
            if (t.exactImage)
                ref_delta(si, t.sites+3*j, t.periodic, t.box, d);
            else
                fast_delta(si, t.sites+3*j, t.periodic, t.box, inv, d);
            hit = ref_norm2(d) < t.cutoff2;
        }
        const unsigned long long mine = (wave_ballot(hit) >> (g*GROUP)) & groupMask;
        const int slot = m+__popcll(mine & ((1ULL << l)-1));
        if (hit) {
            const Real x = Real(d[0]), y = Real(d[1]), z = Real(d[2]);
            const Real u = x*x+y*y+z*z;
            R4 dr = {x, y, z, u};
            sDelta[slot] = dr;
            const Real alpha = Real(t.siteParams[4*j]), w = Real(t.siteParams[4*j+1]);
            R4 pr = {alpha*w/(wI+w), w, wI+w, (t.siteParams[4*j+3] != 0 ? -fast_rcp(u) : fast_rcp(u))};
            sParam[slot] = pr;
            sIndex[slot] = j;
            sMol[slot] = t.molecule[j];
            sWidth[slot] = Real(t.siteParams[4*j+2]);
            sForce[3*slot] = sForce[3*slot+1] = sForce[3*slot+2] = 0;
        }
        if (a < listLength)
            sSlot[a] = (hit ? slot : -1);
        m += __popcll(mine);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const Real cI = t.scaleFactor*Real(1.5)*aI*wI;
    double energy = 0;
    int count = 0;
    // the pairs in rotation order; rot = floor(p / m) from (p + 1/2) / m in FP32 (p < 2016, m <= 64: at least
    // 1/(2 m) from an integer)
    const int P = m*(m-1)/2;
    const float invM = (m > 0 ? 1.0f/(float) m : 0.0f);
    for (int p0 = 0; __any(p0 < P); p0 += GROUP) {
        const int p = p0+l;
        if (p < P) {
            const int rot = (int) (((float) p+0.5f)*invM);
            const int a = p-rot*m;
            int b = a+rot+1;
            if (b >= m)
                b -= m;
            Real e, fa[3], fb[3];
            if (sMol[a] != sMol[b] &&
                plain_triple<Real, DAMPED>(t, sDelta[a], sParam[a], sWidth[a], sIndex[a], sDelta[b], sParam[b], sWidth[b], sIndex[b], cI, sI,
                                           alchI, e, fa, fb)) {
                energy += (double) e;
                count++;
                for (int k = 0; k < 3; k++) {
                    atomicAdd(&sForce[3*a+k], (unsigned long long) to_fixed(fa[k]));
                    atomicAdd(&sForce[3*b+k], (unsigned long long) to_fixed(fb[k]));
                }
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // the neighbour forces as plain stores, one per list entry (zero beyond the cutoff): k_gather_site_forces sums
    // them per site, and the force on A as their negated sum (integer sums: exact in any order)
    for (int a = l; a < listLength; a += GROUP) {
        const int slot = sSlot[a];
        for (int k = 0; k < 3; k++)
            t.contribution[3*((long long) begin+a)+k] = (slot >= 0 ? (long long) sForce[3*slot+k] : 0LL);
    }
    const double e = group_sum<GROUP>(energy);
    const int c = group_sum<GROUP>(count);
    if (valid && l == 0) {
        t.itemEnergy[i] = e;
        t.itemCount[i] = c;
    }
}

// The energy and forces of one triple (A, a, b) of a segmented handle: A the work item's site, a and b two of its
// neighbours (staged vectors AB and AC with their squares, parameters as the staging writes them; b's width and site
// index are read only where they are needed), decided within the cutoff in FP32 (in FP64 near it).
// InteraxThreeBodyTerms.h:51-146, ReferenceInteraxThreeBodyIxn.cpp:104-134 and InteraxThreeBodyForceImpl.h:69-72.
// Returns false for a triple with no contribution (beyond the cutoff or C9 = 0).  Only the forces on a and b are
// formed, as combinations of AB and AC (BC = AC - AB): the three forces of a triple sum to zero, and the force on A
// is taken as the negated sum of its neighbours' fixed-point forces.  The staged parameters of neighbour j are
// (alpha_j w_j / (w_A + w_j), w_j, w_A + w_j, +-1/|Aj|^2), so that C9 = 1.5 c alpha_A w_A P_a P_b (w_A+w_a+w_b) /
// (w_a+w_b) with P_j the first of them.
template <typename Real>
struct TripleResult {
    Real e;
    Real aA, aB;   // the force on a: AB aA + AC aB
    Real bA, bB;   // the force on b: AB bA + AC bB
};

// acc += AB kA + AC kB, as two multiply-adds per component
template <typename Real, typename R4>
__device__ __forceinline__ void add_combination(Real (&acc)[3], const R4& A, const R4& B, Real kA, Real kB) {
    acc[0] = fma(A.x, kA, acc[0]);
    acc[1] = fma(A.y, kA, acc[1]);
    acc[2] = fma(A.z, kA, acc[2]);
    acc[0] = fma(B.x, kB, acc[0]);
    acc[1] = fma(B.y, kB, acc[1]);
    acc[2] = fma(B.z, kB, acc[2]);
}

template <typename Real, bool DAMPED>
__device__ __forceinline__ bool segment_triple(const TripleArgs<Real>& t, const typename Vec4Of<Real>::type& va,
                                               const typename Vec4Of<Real>::type& pa, Real widthA, int siteA,
                                               const typename Vec4Of<Real>::type& vb, const typename Vec4Of<Real>::type& pb,
                                               const Real* sWidth, const int* sIndex, int b, Real cI, Real sI, bool alchI,
                                               TripleResult<Real>& out) {
    Real c[3] = {vb.x-va.x, vb.y-va.y, vb.z-va.z};
    Real w = c[0]*c[0]+c[1]*c[1]+c[2]*c[2];
    bool ok;
    if (fabs(w-t.cutoff2r) < t.band*t.cutoff2r) {
        // segments are nonperiodic: BC from the two centroids in FP64
        double d[3];
        ref_delta(t.sites+3*siteA, t.sites+3*sIndex[b], 0, t.box, d);
        ok = ref_norm2(d) < t.cutoff2;
        c[0] = Real(d[0]);
        c[1] = Real(d[1]);
        c[2] = Real(d[2]);
        w = c[0]*c[0]+c[1]*c[1]+c[2]*c[2];
    }
    else
        ok = w < t.cutoff2r;
    if (!ok)
        return false;
    Real c9 = cI*pa.x*pb.x*(pa.z+pb.y)*fast_rcp(pa.y+pb.y);
    if (t.lambda != Real(1) && (alchI || pa.w < 0 || pb.w < 0))
        c9 *= t.lambda;
    if (c9 == 0)
        return false;
    const Real A[3] = {va.x, va.y, va.z}, B[3] = {vb.x, vb.y, vb.z};
    const Real u = va.w, v = vb.w;
    const Real pp = A[0]*B[0]+A[1]*B[1]+A[2]*B[2];
    const Real q = A[0]*c[0]+A[1]*c[1]+A[2]*c[2];
    const Real s = B[0]*c[0]+B[1]*c[1]+B[2]*c[2];
    const Real uv = u*v;
    const Real rs = fast_rsqrt(uv*w);
    const Real invP = rs*rs;
    const Real invP32 = invP*rs;
    const Real invP52 = invP32*invP;
    const Real pqs = pp*q*s;
    Real e = c9*(invP32-3*pqs*invP52);
    const Real radial = c9*(-3*invP32+15*pqs*invP52);
    const Real cross = -3*c9*invP52;
    const Real ru = radial*fabs(pa.w), rv = radial*fabs(pb.w), rw = radial*(invP*uv);
    const Real X = pp*s*cross, Y = pp*q*cross, Z = q*s*cross;
    Real aA = X-ru, aB = Y-Z, aC = rw-X;
    Real bA = -(Z+X), bB = -(rv+Y), bC = -(Y+rw);
    if constexpr (DAMPED) {
        const Real widthB = sWidth[b];
        Real len[3] = {sqrt(u), sqrt(v), sqrt(w)};
        Real rates[3] = {2*t.dampingScale/(sI+widthA), 2*t.dampingScale/(sI+widthB), 2*t.dampingScale/(widthA+widthB)};
        Real f[3], dfdr[3];
        for (int k = 0; k < 3; k++) {
            if (t.useDamping) {
                tt3<Real>(rates[k]*len[k], f[k], dfdr[k]);
                dfdr[k] *= rates[k];
            }
            else {
                f[k] = 1;
                dfdr[k] = 0;
            }
            if (t.invDampingRadius > 0) {
                Real x = len[k]*t.invDampingRadius;
                Real x6 = x*x*x;
                x6 *= x6;
                Real g = -expm1(-x6);
                Real dg = 6*x6*exp(-x6)/len[k];
                dfdr[k] = dfdr[k]*g + f[k]*dg;
                f[k] *= g;
            }
        }
        const Real damping = f[0]*f[1]*f[2];
        const Real sa = e*dfdr[0]*f[1]*f[2]/len[0];
        const Real sb = e*f[0]*dfdr[1]*f[2]/len[1];
        const Real sc = e*f[0]*f[1]*dfdr[2]/len[2];
        aA = aA*damping-sa;
        aB *= damping;
        aC = aC*damping+sc;
        bA *= damping;
        bB = bB*damping-sb;
        bC = bC*damping-sc;
        e *= damping;
    }
    out.e = e;
    out.aA = aA-aC;
    out.aB = aB+aC;
    out.bA = bA-bC;
    out.bB = bB+bC;
    return true;
}

// Lane l receives the value of lane (l+1) mod 64 (DPP wave_rol:1; every lane active).
__device__ __forceinline__ float lane_rotate(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x134, 0xF, 0xF, false));
}
__device__ __forceinline__ double lane_rotate(double v) {
    return wave_rotate_down(v);
}

// Sum over the 16 lanes of each row, every lane of the row receiving the same order-fixed total.
__device__ __forceinline__ float row_sum(float v) {
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, false));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, false));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x141, 0xF, 0xF, false));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x140, 0xF, 0xF, false));
    return v;
}
__device__ __forceinline__ double row_sum(double v) {
    for (int offset = 1; offset < 16; offset <<= 1)
        v += __shfl_xor(v, offset, 64);
    return v;
}

// Steps (64 pairs each) of the triple kernel of segmented handles for a site with m neighbours: every group but the
// last has 32 rotation steps and m - 64 (k + 1) later-block steps, the last group the rotation steps of its nb
// neighbours.
__host__ __device__ __forceinline__ long long segment_steps(long long m) {
    if (m < 2)
        return 0;
    const long long last = (m-1)/64, nb = m-64*last;
    return last*m-32*last*last+(nb*(nb-1)/2+63)/64;
}

// 3b. Segmented handles (clusters): one workgroup per work item, a site A and consecutive shares of its neighbour
// pairs, one wavefront per share (blockDim.x / 64 wavefronts take every so-many-th share).  The pairs a < b of the
// site's m neighbours come in steps of 64 (one per lane), grouped by blocks of 64 neighbours: group k holds first the
// pairs within block k (neighbours 64 k .. 64 k + nb - 1) in rotation order -- rotation r holds (x, (x+r+1) mod nb)
// -- then the pairs of block k with each later block l, 64 steps for a full block l (step s: lane x holds
// a = 64 k + x and b = 64 l + (x+s) mod 64) and one step per neighbour b of a last, partial block (lane x holds
// a = 64 k + x and the one b).  Group k starts at step k m - 32 k^2.
//   In a full block k lane x holds a = 64 k + x in every step of the group: its loads are made once and the force on
// a is summed in the lane's registers over the group.  In the rotation steps of a full block and in the steps with a
// full block l, the b of lane x is the previous step's b of lane x+1, so the force on b is summed in a register that
// moves down one lane per step (DPP wave rotation) and is added to LDS in fixed point once per block pair; with a
// partial block l the one b's force is summed over each row of 16 lanes, and in the rotation steps of a partial
// block every lane adds its own a and b.  The force on A is the negated sum of its neighbours' forces (fixed point,
// exact).  The shares depend on the site's list capacity only, so the results do not depend on the launch shape (one
// instantiation for every number of wavefronts).
template <typename Real, bool DAMPED>
__global__ void __launch_bounds__(512) k_triple_groups(TripleArgs<Real> t) {
    typedef typename Vec4Of<Real>::type R4;
    extern __shared__ double sharedStorage[];
    R4* sDelta = (R4*) sharedStorage;
    R4* sParam = sDelta+t.capacity;
    unsigned long long* sForce = (unsigned long long*) (sParam+t.capacity);
    int* sIndex = (int*) (sForce+3*(size_t) t.capacity);
    int* sMol = sIndex+t.capacity;
    Real* sWidth = (Real*) (sMol+t.capacity);
    const int item = (t.itemIndex != nullptr ? t.itemIndex[blockIdx.x] : (int) blockIdx.x);
    const int i = t.itemSite[item];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6, waves = (int) blockDim.x >> 6;
    const int begin = t.start[i], m = t.count[i];
    const Real aI = Real(t.siteParams[4*i]), wI = Real(t.siteParams[4*i+1]), sI = Real(t.siteParams[4*i+2]);
    const bool alchI = t.siteParams[4*i+3] != 0;
    // the work item's forces on the neighbours, summed over 16 slots per component (few lanes per LDS address)
    __shared__ unsigned long long sSiteSum[3][16];
    if (threadIdx.x < 48)
        sSiteSum[threadIdx.x >> 4][threadIdx.x & 15] = 0;
    const double* si = t.sites+3*i;
    for (int a = threadIdx.x; a < m; a += blockDim.x) {
        // AB from the two centroids, as the neighbour search rounds it (segments are nonperiodic)
        const int j = t.nbrIndex[begin+a];
        const Real x = Real(__dsub_rn(t.sites[3*j], si[0])), y = Real(__dsub_rn(t.sites[3*j+1], si[1])),
                   z = Real(__dsub_rn(t.sites[3*j+2], si[2]));
        const Real u = x*x+y*y+z*z;
        R4 dr = {x, y, z, u};
        sDelta[a] = dr;
        const Real alpha = Real(t.siteParams[4*j]), w = Real(t.siteParams[4*j+1]);
        R4 pr = {alpha*w/(wI+w), w, wI+w, (t.siteParams[4*j+3] != 0 ? -fast_rcp(u) : fast_rcp(u))};
        sParam[a] = pr;
        sIndex[a] = j;
        sMol[a] = t.molecule[j];
        sWidth[a] = Real(t.siteParams[4*j+2]);
        sForce[3*a] = sForce[3*a+1] = sForce[3*a+2] = 0;
    }
    __syncthreads();
    const Real cI = t.scaleFactor*Real(1.5)*aI*wI;
    const int2 sh = t.itemShares[item];
    const int numShares = sh.x, shareSteps = sh.y/64;
    const int firstStep = t.itemFirst[item]/64;
    const int K = (m+63)/64;
    const int total = (int) segment_steps(m);
    // add a register sum of forces to neighbour b in fixed point
    auto addForce = [&](int b, const Real (&f)[3]) {
        for (int c = 0; c < 3; c++) {
            const long long v = to_fixed(f[c]);
            if (v != 0)
                atomicAdd(&sForce[3*b+c], (unsigned long long) v);
        }
    };
    for (int v = wave; v < numShares; v += waves) {
        const int g0 = firstStep+v*shareSteps, g1 = min(total, g0+shareSteps);
        double energy = 0;
        int count = 0;   // the share's triples (wavefront-uniform)
        int g = g0;
        // the group of step g0 and the step's place in it
        int k = 0, groupStart = 0;
        while (k < K-1 && g0 >= groupStart+32+(m-64*(k+1))) {
            groupStart += 32+(m-64*(k+1));
            k++;
        }
        int local = g0-groupStart;
        while (g < g1) {
            const int nb = min(64, m-64*k);
            const bool full = (nb == 64);
            const int rotSteps = (nb*(nb-1)/2+63)/64;
            const int groupSteps = rotSteps+max(0, m-64*(k+1));
            // the lane's own neighbour a = 64 k + lane (full blocks)
            R4 oDelta = {0, 0, 0, 0}, oParam = {0, 0, 0, 0};
            Real oWidth = 0;
            int oMol = 0, oSite = 0;
            if (full) {
                const int a = 64*k+lane;
                oDelta = sDelta[a];
                oParam = sParam[a];
                oWidth = sWidth[a];
                oMol = sMol[a];
                oSite = sIndex[a];
            }
            Real accA[3] = {0, 0, 0};
            while (g < g1 && local < groupSteps) {
                TripleResult<Real> r;
                if (local < rotSteps && full) {
                    // rotation steps of a full block: lane x holds (x, (x+rot+1) mod 64), rotation 31 only x < 32
                    const int r0 = local, r1 = min(rotSteps, local+(g1-g));
                    Real accB[3] = {0, 0, 0};
                    for (int rot = r0; rot < r1; rot++) {
                        const int b = 64*k+((lane+rot+1) & 63);
                        const R4 vb = sDelta[b];
                        bool hit = false;
                        if ((rot < 31 || lane < 32) && oMol != sMol[b] &&
                            segment_triple<Real, DAMPED>(t, oDelta, oParam, oWidth, oSite, vb, sParam[b], sWidth, sIndex, b, cI, sI, alchI, r)) {
                            hit = true;
                            energy += (double) r.e;
                            add_combination(accA, oDelta, vb, r.aA, r.aB);
                            add_combination(accB, oDelta, vb, r.bA, r.bB);
                        }
                        count += __popcll(wave_ballot(hit));
                        for (int c = 0; c < 3; c++)
                            accB[c] = lane_rotate(accB[c]);
                    }
                    addForce(64*k+((lane+r1+1) & 63), accB);
                    g += r1-r0;
                    local = r1;
                }
                else if (local < rotSteps) {
                    // rotation steps of a partial block: pair p of its nb (nb - 1) / 2; rot = floor(p / nb) from
                    // (p + 1/2) / nb in FP32, which lies at least 1/(2 nb) from an integer (p < 2016, nb < 64)
                    const int r1 = min(rotSteps, local+(g1-g));
                    const float invNb = 1.0f/(float) nb;
                    for (int step = local; step < r1; step++) {
                        const int p = 64*step+lane;
                        bool hit = false;
                        if (p < nb*(nb-1)/2) {
                            const int rot = (int) (((float) p+0.5f)*invNb);
                            const int x = p-rot*nb;
                            int y = x+rot+1;
                            if (y >= nb)
                                y -= nb;
                            const int a = 64*k+x, b = 64*k+y;
                            const R4 va = sDelta[a], vb = sDelta[b];
                            if (sMol[a] != sMol[b] &&
                                segment_triple<Real, DAMPED>(t, va, sParam[a], sWidth[a], sIndex[a], vb, sParam[b], sWidth, sIndex, b, cI, sI, alchI, r)) {
                                hit = true;
                                energy += (double) r.e;
                                Real fa[3] = {0, 0, 0}, fb[3] = {0, 0, 0};
                                add_combination(fa, va, vb, r.aA, r.aB);
                                add_combination(fb, va, vb, r.bA, r.bB);
                                addForce(a, fa);
                                addForce(b, fb);
                            }
                        }
                        count += __popcll(wave_ballot(hit));
                    }
                    g += r1-local;
                    local = r1;
                }
                else {
                    const int t0 = local-rotSteps;
                    const int l = k+1+(t0 >> 6), s0 = t0 & 63;
                    const int nbl = min(64, m-64*l);
                    const int s1 = min(nbl, s0+(g1-g));
                    if (nbl == 64) {
                        // block pair (k, l): lane x holds a = 64 k + x and b = 64 l + (x+s) mod 64
                        Real accB[3] = {0, 0, 0};
                        for (int step = s0; step < s1; step++) {
                            const int b = 64*l+((lane+step) & 63);
                            const R4 vb = sDelta[b];
                            bool hit = false;
                            if (oMol != sMol[b] &&
                                segment_triple<Real, DAMPED>(t, oDelta, oParam, oWidth, oSite, vb, sParam[b], sWidth, sIndex, b, cI, sI, alchI, r)) {
                                hit = true;
                                energy += (double) r.e;
                                add_combination(accA, oDelta, vb, r.aA, r.aB);
                                add_combination(accB, oDelta, vb, r.bA, r.bB);
                            }
                            count += __popcll(wave_ballot(hit));
                            for (int c = 0; c < 3; c++)
                                accB[c] = lane_rotate(accB[c]);
                        }
                        addForce(64*l+((lane+s1) & 63), accB);
                    }
                    else {
                        // a partial block l: one b per step, its force summed over each row of 16 lanes; lane
                        // (step mod 16) of each row keeps the row's sum, added to LDS for 16 steps at a time
                        Real keep[3] = {0, 0, 0};
                        for (int step = s0; step < s1; step++) {
                            const int b = 64*l+step;
                            const R4 vb = sDelta[b];
                            Real fb[3] = {0, 0, 0};
                            bool hit = false;
                            if (oMol != sMol[b] &&
                                segment_triple<Real, DAMPED>(t, oDelta, oParam, oWidth, oSite, vb, sParam[b], sWidth, sIndex, b, cI, sI, alchI, r)) {
                                hit = true;
                                energy += (double) r.e;
                                add_combination(accA, oDelta, vb, r.aA, r.aB);
                                add_combination(fb, oDelta, vb, r.bA, r.bB);
                            }
                            count += __popcll(wave_ballot(hit));
                            const bool slot = ((lane & 15) == (step & 15));
                            for (int c = 0; c < 3; c++) {
                                const Real sum = row_sum(fb[c]);
                                if (slot)
                                    keep[c] = sum;
                            }
                            if ((step & 15) == 15 || step == s1-1) {
                                // lanes of slots without a step of this window keep zero and add nothing
                                addForce(64*l+(step & ~15)+(lane & 15), keep);
                                for (int c = 0; c < 3; c++)
                                    keep[c] = 0;
                            }
                        }
                    }
                    g += s1-s0;
                    local += s1-s0;
                }
            }
            if (full)
                addForce(64*k+lane, accA);
            if (local == groupSteps) {
                k++;
                local = 0;
            }
        }
        // the share's energy and triple count
        const double e = wave_sum(energy);
        if (lane == 0) {
            t.itemEnergy[t.itemShareBase[item]+v] = e;
            t.itemCount[t.itemShareBase[item]+v] = count;
        }
    }
    __syncthreads();
    // the neighbour forces (integer sums: the order of the atomics does not matter) and, negated, their sum on A
    long long sum[3] = {0, 0, 0};
    for (int a = threadIdx.x; a < m; a += blockDim.x) {
        const int j = sIndex[a];
        for (int c = 0; c < 3; c++) {
            const long long f = (long long) sForce[3*a+c];
            add_fixed(&t.siteForce[(size_t) c*t.numSites+j], f);
            sum[c] += f;
        }
    }
    for (int c = 0; c < 3; c++)
        if (sum[c] != 0)
            atomicAdd(&sSiteSum[c][lane & 15], (unsigned long long) sum[c]);
    __syncthreads();
    if (threadIdx.x < 3) {
        unsigned long long total = 0;
        for (int k = 0; k < 16; k++)
            total += sSiteSum[threadIdx.x][k];
        add_fixed(&t.siteForce[(size_t) threadIdx.x*t.numSites+i], -(long long) total);
    }
}

// Segmented neighbour lists: site i's candidates are the later sites of its own segment, [i+1, siteEnd[i]); the hits
// (another molecule, within the cutoff, no periodic images) are written from start[i], their number to count[i].
constexpr int SEGMENT_SEARCH_WAVES = 4;   // sites per workgroup of k_neighbors_segments, one wavefront each
__global__ void __launch_bounds__(64*SEGMENT_SEARCH_WAVES) k_neighbors_segments(const double* sites, const int* molecule, const int* siteEnd,
                                                                             double cutoff2, const int* start, int* count, int* nbrIndex,
                                                                             const int* siteIndex, int numSearched) {
    const int w = SEGMENT_SEARCH_WAVES*(int) blockIdx.x+(int) (threadIdx.x >> 6);
    if (w >= numSearched)
        return;   // a whole wavefront
    const int i = (siteIndex != nullptr ? siteIndex[w] : w);
    const int lane = threadIdx.x & 63;
    const double* si = sites+3*i;
    const int mi = molecule[i];
    const int end = siteEnd[i];
    int written = 0;
    for (int j0 = i+1; j0 < end; j0 += 64) {
        const int j = j0+lane;
        bool hit = false;
        double d[3] = {0, 0, 0};
        if (j < end && molecule[j] != mi) {
            d[0] = __dsub_rn(sites[3*j], si[0]);
            d[1] = __dsub_rn(sites[3*j+1], si[1]);
            d[2] = __dsub_rn(sites[3*j+2], si[2]);
            hit = ref_norm2(d) < cutoff2;
        }
        const unsigned long long ballot = __ballot(hit);
        if (hit)
            nbrIndex[start[i]+written+__popcll(ballot & ((1ULL << lane)-1))] = j;
        written += __popcll(ballot);
    }
    if (lane == 0)
        count[i] = written;
}

// Energy and triple count of every segment: those of its shares (ranges[s] .. ranges[s+1]) summed in order (one
// workgroup per segment, fixed tree), into segEnergy[s] and segCount[s], and the energy added to energies[s] (if not
// NULL).
__global__ void __launch_bounds__(256) k_reduce_site_segments(const double* shareEnergy, const int* shareCount, const int* ranges,
                                                               const int* segList, double* segEnergy, int* segCount,
                                                               double* energies) {
    __shared__ double sE[256];
    __shared__ int sC[256];
    const int seg = (segList != nullptr ? segList[blockIdx.x] : (int) blockIdx.x);
    double e = 0;
    int c = 0;
    for (int k = ranges[seg]+threadIdx.x; k < ranges[seg+1]; k += 256) {
        e += shareEnergy[k];
        c += shareCount[k];
    }
    sE[threadIdx.x] = e;
    sC[threadIdx.x] = c;
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) {
            sE[threadIdx.x] += sE[threadIdx.x+stride];
            sC[threadIdx.x] += sC[threadIdx.x+stride];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        segEnergy[seg] = sE[0];
        segCount[seg] = sC[0];
        if (energies != nullptr)
            energies[seg] += sE[0];
    }
}

__global__ void k_zero_long(long long* p, long long n) {
    long long i = (long long) blockIdx.x*blockDim.x+threadIdx.x;
    if (i < n)
        p[i] = 0;
}

// 4. ReferenceInteraxThreeBodyIxn.cpp:129-133
__global__ void k_apply_site_forces(const long long* siteForce, int numSites, const int* offsets, const int* particles,
                                    long long* force, int stride) {
    int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= numSites)
        return;
    double scale = 1.0/(double) (offsets[i+1]-offsets[i]);
    long long f[3];
    for (int k = 0; k < 3; k++) {
        double v = (double) siteForce[(size_t) k*numSites+i]/FIXED_POINT_SCALE;
        f[k] = (long long) (v*scale*FIXED_POINT_SCALE);
    }
    for (int p = offsets[i]; p < offsets[i+1]; p++) {
        int particle = particles[p];
        for (int k = 0; k < 3; k++)
            add_fixed(&force[(size_t) k*stride+particle], f[k]);
    }
}

// Plain handles: the force on every site, the sum of its own part (the triples it owns) and of the list entries that
// name it (its reverse list, built with the lists), shared equally among its particles as k_apply_site_forces does.
// GATHER_LANES lanes per site take every GATHER_LANES-th entry, so that the loads of a site are in flight together;
// their integer sums are combined by lane exchanges (exact in any order).  Each workgroup also sums the energies and
// triple counts of its sites in site order into one partial (groupEnergy, groupCount), for k_reduce_sites.
constexpr int GATHER_LANES = 16;
constexpr int GATHER_THREADS = 256;
__global__ void __launch_bounds__(GATHER_THREADS) k_gather_site_forces(const long long* own, const int* listStart,
                                                                       const int* listCount, const long long* contribution,
                                                                       const int* revStart, const int* revEnd,
                                                                       const int* revList, int numSites, const int* offsets,
                                                                       const int* particles, long long* force, int stride,
                                                                       int applyForces, const double* siteEnergy,
                                                                       const int* siteCount, double* groupEnergy,
                                                                       int* groupCount) {
    const int i = (int) ((blockIdx.x*blockDim.x+threadIdx.x)/GATHER_LANES);
    const int sub = threadIdx.x % GATHER_LANES;
    {
        constexpr int SITES = GATHER_THREADS/GATHER_LANES;
        __shared__ double sE[SITES];
        __shared__ int sC[SITES];
        if (threadIdx.x < SITES) {
            const int site = blockIdx.x*SITES+threadIdx.x;
            sE[threadIdx.x] = (site < numSites ? siteEnergy[site] : 0.0);
            sC[threadIdx.x] = (site < numSites ? siteCount[site] : 0);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double e = 0;
            int c = 0;
            for (int k = 0; k < SITES; k++) {
                e += sE[k];
                c += sC[k];
            }
            groupEnergy[blockIdx.x] = e;
            groupCount[blockIdx.x] = c;
        }
    }
    if (!applyForces)
        return;   // uniform over the launch: the lane exchanges below need every lane of the wave
    long long total[3] = {0, 0, 0};
    if (i < numSites) {
        if (listStart == nullptr) {
            if (sub == 0)
                for (int k = 0; k < 3; k++)
                    total[k] = own[(size_t) k*numSites+i];
        }
        else
            // the force on site i from its own triples: the negated sum of its list's contributions
            for (int a = sub; a < listCount[i]; a += GATHER_LANES) {
                const long long e = (long long) listStart[i]+a;
                for (int k = 0; k < 3; k++)
                    total[k] -= contribution[3*e+k];
            }
        for (int r = revStart[i]+sub; r < revEnd[i]; r += GATHER_LANES) {
            const long long e = revList[r];
            for (int k = 0; k < 3; k++)
                total[k] += contribution[3*e+k];
        }
    }
    for (int offset = GATHER_LANES/2; offset > 0; offset >>= 1)
        for (int k = 0; k < 3; k++)
            total[k] += __shfl_xor(total[k], offset, 64);
    if (i >= numSites || sub != 0)
        return;
    const double scale = 1.0/(double) (offsets[i+1]-offsets[i]);
    long long f[3];
    for (int k = 0; k < 3; k++)
        f[k] = (long long) ((double) total[k]/FIXED_POINT_SCALE*scale*FIXED_POINT_SCALE);
    for (int p = offsets[i]; p < offsets[i+1]; p++) {
        const int particle = particles[p];
        for (int k = 0; k < 3; k++)
            add_fixed(&force[(size_t) k*stride+particle], f[k]);
    }
}

// Reverse lists of the neighbour lists (list entry e = i capacity + a of site i names site j): count per site (pass 0),
// then, after a scan of the counts (k_cell_scan), the entries of each site in any order (pass 1; the sums over them
// are integer, so their order does not matter).
__global__ void k_reverse_lists(const int* nbrIndex, const int* count, int numSites, int capacity, int* revCount, int* revCursor,
                                int* revList) {
    const long long e = (long long) blockIdx.x*blockDim.x+threadIdx.x;
    if (e >= (long long) numSites*capacity)
        return;
    const int i = (int) (e/capacity), a = (int) (e-(long long) i*capacity);
    if (a >= count[i])
        return;
    const int j = nbrIndex[e];
    if (revList == nullptr)
        atomicAdd(&revCount[j], 1);
    else
        revList[atomicAdd(&revCursor[j], 1)] = (int) e;
}

// The total energy and triple count of a launch: one workgroup of 1024 threads, strided (coalesced) partial sums
// with four loads in flight per thread, then a fixed tree, so the result does not depend on timing.
__global__ void __launch_bounds__(1024) k_reduce_sites(const double* siteEnergy, const int* siteCount, int numSites, double* energy,
                                                       long long* tripleCount, int addEnergy) {
    __shared__ double sE[1024];
    __shared__ long long sC[1024];
    double e = 0;
    long long c = 0;
    for (int k0 = threadIdx.x; k0 < numSites; k0 += 4*1024) {
        double v[4] = {0, 0, 0, 0};
        int n[4] = {0, 0, 0, 0};
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (k0+u*1024 < numSites) {
                v[u] = siteEnergy[k0+u*1024];
                n[u] = siteCount[k0+u*1024];
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
        *tripleCount = sC[0];
    }
}

int grid_of(long long n, int threads) {
    return (int) std::max(1LL, (n+threads-1)/threads);
}

}  // namespace

struct ixh_threebody {
    int device = 0;
    int numParticles = 0;
    int numSites = 0;
    int precision = IXH_MIXED;
    double cutoff = 1.2;
    double scaleFactor = 1, dampingScale = 1, centroidDampingRadius = 0;
    int dampingMethod = 0;
    int periodic = 0;
    bool anyAlchemical = false;
    std::vector<int> offsets, particles, molecule;
    std::vector<unsigned char> alchemical;
    std::vector<double> params;   // alpha, frequency, width, alchemical
    DeviceArray<int> offsetsDev, particlesDev, moleculeDev, start, count, nbrIndex, siteCount;
    // cell grid of the neighbour search
    DeviceArray<unsigned> cellKeys, cellKeysSorted;
    DeviceArray<int> cellValues, sortedSites, cellStart, cellEnd;
    DeviceArray<int> cellCount, cellCursor;   // counting sort by cell (cellCount zero between evaluations)
    DeviceArray<char> sortTemp;
    DeviceArray<double> bbox;
    // list totals of the neighbour search read back by the host: the number of entries and the longest list
    DeviceArray<int> listTotals;
    int* hostTotals = nullptr;   // pinned, 2 ints
    int listCapacity = 0;        // entries per site of the unsegmented lists
    int listSites = -1;          // sites the lists were sized for
    // plain handles: lists within cutoff + skin, rebuilt when a site has moved half the skin since buildSites
    bool listBuilt = false;
    int listMax = 0;             // the longest list of the last build (sizes the triple kernel's LDS)
    DeviceArray<double> buildSites;
    int* hostFlag = nullptr;     // pinned, coherent: the list-check flags of the two frames, written by the check kernel
    int frame = 0;               // the frame of the last check
    bool deferredCheck = false;  // the last call checked without waiting (IXH_DEFER_LIST_CHECK)
    hipEvent_t checked = nullptr;   // the end of the last deferred check
    // per-entry neighbour forces and the reverse lists that sum them per site (plain handles)
    DeviceArray<long long> contribution;
    DeviceArray<int> revCount, revStart, revEnd, revCursor, revList;
    DeviceArray<double> groupEnergy;   // energies and triple counts per workgroup of k_gather_site_forces
    DeviceArray<int> groupCount;
    // independent systems (segments) of a batch: sites [segOffsets[s], segOffsets[s+1]); neighbour lists of static
    // capacity (the later sites of the same segment), so that an evaluation needs no host synchronization
    int numSegments = 0;
    std::vector<int> segOffsets;
    DeviceArray<int> segOffsetsDev, siteEnd;
    DeviceArray<double> segEnergy;   // energy and triple count of every segment of the last evaluation
    DeviceArray<int> segCount;
    int maxSegment = 0;
    // work items of the triple kernel (segmented handles): a site and consecutive shares of its neighbour pairs, items
    // of segment s in [segItems[s], segItems[s+1]); shares of segment s in [segShares[s], segShares[s+1])
    int numItems = 0, numShares = 0;
    std::vector<int> segItems, segShares;
    // launch shape of the triple kernel of segmented handles: wavefronts per work item, neighbour pairs per work item
    int tripleWaves = 4;
    int chunkPairs = TRIPLE_CHUNK;
    DeviceArray<int> itemSite, itemShareBase, segItemsDev, segSharesDev, subsetItems;
    DeviceArray<int> itemFirst;
    DeviceArray<int2> itemShares;
    DeviceArray<int> itemOrder;   // all items, most pairs first (the launch order of a full evaluation)
    // subsets of segments: site list and segment list of the last call
    StagingRing subsetStaging;
    DeviceArray<int> subsetSites, subsetSegments;
    // plain handles with lists of at most 64 entries: the sites by list length, longest first (the packed triple
    // kernel's order, so that the sites of a wavefront have similar work); valid after a list build
    DeviceArray<int> lengthOrder, lengthKeys, siteIota;
    bool lengthOrderValid = false;
    DeviceArray<double> paramsDev, sites, siteEnergy;
    DeviceArray<long long> siteForce, tripleCount;
    hipStream_t lastStream = nullptr;
    hipEvent_t done = nullptr;   // recorded at the end of every execute: calls on other streams wait for it
    hipEvent_t searched = nullptr;   // segmented handles: recorded after the neighbour search, if set (not owned)
    // parameter updates: FP64 alpha, frequency, width of the host path, the update's end, and its width check
    DeviceArray<double> rawSites;
    DeviceArray<int> invalid;
    int* hostInvalid = nullptr;   // pinned
    hipEvent_t updateDone = nullptr;
    bool updateRecorded = false, updatePending = false;
    KernelTimer timer;
};

namespace {

// Shares of a site's neighbour pairs (segmented handles): about a quarter of the steps of 64 pairs its list can hold
// (segment_steps), between SHARE_MIN and SHARE_MAX pairs.  They depend on the capacity alone, never on the launch shape.
constexpr long long SHARE_MIN = 64, SHARE_MAX = 4096;

// The work items of a segmented handle's triple kernel: every site's shares, about chunkPairs pairs of them per item
// (at least one share), launched with the items of the most pairs first.
void build_items(ixh_threebody* h) {
    std::vector<int> itemSite, shareBase, itemFirst;
    std::vector<long long> work;
    std::vector<int2> itemShares;
    h->segItems.assign(h->numSegments+1, 0);
    h->segShares.assign(h->numSegments+1, 0);
    long long shares = 0;
    for (int g = 0; g < h->numSegments; g++) {
        for (int i = h->segOffsets[g]; i < h->segOffsets[g+1]; i++) {
            const long long cap = h->segOffsets[g+1]-i-1;
            const long long pairs = 64*segment_steps(cap);
            if (pairs > INT32_MAX)
                throw Error(IXH_ERROR_CAPACITY, "ixh_threebody_create_segmented: a segment has too many sites for the triple kernel");
            const long long size = std::min(SHARE_MAX, std::max(SHARE_MIN, ((pairs+3)/4+63)/64*64));
            const long long count = std::max(1LL, (pairs+size-1)/size);
            const long long perItem = std::max(1LL, (long long) h->chunkPairs/size);
            for (long long f = 0; f < count; f += perItem) {
                const long long n = std::min(perItem, count-f);
                itemSite.push_back(i);
                itemFirst.push_back((int) (f*size));
                itemShares.push_back(make_int2((int) n, (int) size));
                shareBase.push_back((int) (shares+f));
                work.push_back(std::max(0LL, std::min(pairs, (f+n)*size)-f*size));
            }
            shares += count;
            if (shares > INT32_MAX || itemSite.size() > (size_t) INT32_MAX)
                throw Error(IXH_ERROR_CAPACITY, "ixh_threebody_create_segmented: too many work items");
        }
        h->segItems[g+1] = (int) itemSite.size();
        h->segShares[g+1] = (int) shares;
    }
    h->numItems = (int) itemSite.size();
    h->numShares = (int) shares;
    std::vector<int> order(itemSite.size());
    for (size_t w = 0; w < order.size(); w++)
        order[w] = (int) w;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
    const size_t items = std::max<size_t>(order.size(), 1);
    h->itemOrder.reserve(items);
    h->itemSite.reserve(items);
    h->itemShareBase.reserve(items);
    h->itemFirst.reserve(items);
    h->itemShares.reserve(items);
    if (!order.empty()) {
        h->itemOrder.upload(order.data(), order.size(), nullptr);
        h->itemSite.upload(itemSite.data(), itemSite.size(), nullptr);
        h->itemShareBase.upload(shareBase.data(), shareBase.size(), nullptr);
        h->itemFirst.upload(itemFirst.data(), itemFirst.size(), nullptr);
        h->itemShares.upload(itemShares.data(), itemShares.size(), nullptr);
    }
    h->segItemsDev.upload(h->segItems.data(), h->segItems.size(), nullptr);
    h->segSharesDev.upload(h->segShares.data(), h->segShares.size(), nullptr);
    h->siteEnergy.reserve(std::max(h->numShares, 1));
    h->siteCount.reserve(std::max(h->numShares, 1));
    IXH_HIP_CHECK(hipStreamSynchronize(nullptr));
}

void check_threebody(const ixh_threebody_desc* d) {
    if (d == nullptr)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody: null descriptor");
    if (d->num_sites < 0 || d->num_particles < 0)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody: negative size");
    if (d->precision != IXH_SINGLE && d->precision != IXH_MIXED && d->precision != IXH_DOUBLE)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody: illegal precision");
    // InteraxThreeBodyForce.cpp:98-101, 115-118 and InteraxThreeBodyForceImpl.cpp:42-59
    if (d->damping_method != IXH_DAMPING_NONE && d->damping_method != IXH_DAMPING_TANG_TOENNIES)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "InteraxThreeBodyForce: Illegal value for damping method");
    if (d->centroid_damping_radius < 0)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "InteraxThreeBodyForce: The centroid damping radius cannot be negative");
    for (int i = 0; i < d->num_sites; i++) {
        if (d->site_offsets[i+1] <= d->site_offsets[i])
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "InteraxThreeBodyForce: A site must contain at least one particle");
        for (int k = d->site_offsets[i]; k < d->site_offsets[i+1]; k++)
            if (d->site_particles[k] < 0 || d->site_particles[k] >= d->num_particles)
                throw Error(IXH_ERROR_INVALID_ARGUMENT, "InteraxThreeBodyForce: Illegal particle index for a site");
        if (d->damping_method == IXH_DAMPING_TANG_TOENNIES && d->width[i] <= 0)
            throw Error(IXH_ERROR_INVALID_ARGUMENT,
                        "InteraxThreeBodyForce: The Tang-Toennies damping requires a positive width for every site");
    }
    if (d->cutoff <= 0)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "InteraxThreeBodyForce: The cutoff distance must be positive");
}

// alpha, frequency and width of every site into the kernels' parameter array (the alchemical flag is kept); a width
// the Tang-Toennies damping cannot take sets *invalid.
__global__ void k_site_params(const double* alpha, const double* frequency, const double* width, int n, int checkWidth,
                              double* params, int* invalid) {
    const int i = blockIdx.x*blockDim.x+threadIdx.x;
    if (i >= n)
        return;
    params[4*(size_t) i] = alpha[i];
    params[4*(size_t) i+1] = frequency[i];
    params[4*(size_t) i+2] = width[i];
    if (checkWidth && !(width[i] > 0))
        *invalid = 1;
}

__global__ void k_clear_flag(int* p) {
    *p = 0;
}

// The device work of a parameter update from FP64 arrays in device memory, on `stream`, after the evaluations and
// the update before it.
void update_site_params(ixh_threebody* h, const double* alpha, const double* frequency, const double* width, hipStream_t stream) {
    if (h->done != nullptr)
        IXH_HIP_CHECK(hipStreamWaitEvent(stream, h->done, 0));
    if (h->updateRecorded)
        IXH_HIP_CHECK(hipStreamWaitEvent(stream, h->updateDone, 0));
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, stream, h->invalid.ptr);
    IXH_HIP_CHECK(hipGetLastError());
    if (h->numSites > 0) {
        hipLaunchKernelGGL(k_site_params, dim3((h->numSites+255)/256), dim3(256), 0, stream, alpha, frequency, width, h->numSites,
                           (int) (h->dampingMethod == IXH_DAMPING_TANG_TOENNIES), h->paramsDev.ptr, h->invalid.ptr);
        IXH_HIP_CHECK(hipGetLastError());
    }
    IXH_HIP_CHECK(hipMemcpyAsync(h->hostInvalid, h->invalid.ptr, sizeof(int), hipMemcpyDeviceToHost, stream));
    IXH_HIP_CHECK(hipEventRecord(h->updateDone, stream));
    h->updateRecorded = true;
    h->updatePending = true;
}

// Before an evaluation on `stream`: it is ordered after the last update, whose widths the Tang-Toennies damping has
// checked on the device (the host waits for that check).
void finish_site_params(ixh_threebody* h, hipStream_t stream) {
    if (!h->updatePending)
        return;
    if (h->dampingMethod == IXH_DAMPING_TANG_TOENNIES) {
        IXH_HIP_CHECK(hipEventSynchronize(h->updateDone));
        if (*h->hostInvalid) {
            h->updatePending = false;
            throw Error(IXH_ERROR_INVALID_ARGUMENT,
                        "InteraxThreeBodyForce: The Tang-Toennies damping requires a positive width for every site");
        }
    }
    IXH_HIP_CHECK(hipStreamWaitEvent(stream, h->updateDone, 0));
    h->updatePending = false;
}

void load_threebody_params(ixh_threebody* h, const ixh_threebody_desc* d) {
    h->params.resize(4*(size_t) h->numSites);
    for (int i = 0; i < h->numSites; i++) {
        h->params[4*i] = d->alpha[i];
        h->params[4*i+1] = d->frequency[i];
        h->params[4*i+2] = d->width[i];
        h->params[4*i+3] = h->alchemical[i];
    }
    h->scaleFactor = d->scale_factor;
    h->dampingScale = d->damping_scale;
    h->centroidDampingRadius = d->centroid_damping_radius;
    h->paramsDev.upload(h->params.data(), h->params.size(), nullptr);
    IXH_HIP_CHECK(hipStreamSynchronize(nullptr));
}

__global__ void k_site_bounds(const double* sites, int numSites, double* out) {
    __shared__ double lo[3][256], hi[3][256];
    double l[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, u[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int i = threadIdx.x; i < numSites; i += blockDim.x)
        for (int k = 0; k < 3; k++) {
            l[k] = fmin(l[k], sites[3*i+k]);
            u[k] = fmax(u[k], sites[3*i+k]);
        }
    for (int k = 0; k < 3; k++) {
        lo[k][threadIdx.x] = l[k];
        hi[k][threadIdx.x] = u[k];
    }
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride)
            for (int k = 0; k < 3; k++) {
                lo[k][threadIdx.x] = fmin(lo[k][threadIdx.x], lo[k][threadIdx.x+stride]);
                hi[k][threadIdx.x] = fmax(hi[k][threadIdx.x], hi[k][threadIdx.x+stride]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) {
            out[k] = lo[k][0];
            out[3+k] = hi[k][0];
        }
}

// The cell grid of one evaluation.  Periodic boxes: rectangular boxes with at least three cells of the cutoff per
// axis; other boxes, and axes with fewer cells, get one cell along that axis.  Nonperiodic systems of more than
// 4096 sites: a grid of cutoff-sized cells over the sites' bounding box (at most 64 per axis).
CellGrid make_grid(ixh_threebody* h, const double* box, double cellSize, hipStream_t stream) {
    CellGrid g;
    g.periodic = h->periodic;
    for (int k = 0; k < 3; k++) {
        g.n[k] = 1;
        g.origin[k] = 0;
        g.scale[k] = 0;
    }
    const double cell = cellSize*(1+1e-12);
    if (h->periodic) {
        bool rectangular = (box[3] == 0 && box[6] == 0 && box[7] == 0);
        if (rectangular)
            for (int k = 0; k < 3; k++) {
                double L = box[4*k];
                int cells = (int) std::min(1024.0, std::floor(L/cell));
                if (cells >= 3) {
                    g.n[k] = cells;
                    g.scale[k] = 1.0/L;
                }
            }
    }
    else if (h->numSites > 4096) {
        h->bbox.reserve(6);
        hipLaunchKernelGGL(k_site_bounds, dim3(1), dim3(256), 0, stream, h->sites.ptr, h->numSites, h->bbox.ptr);
        IXH_HIP_CHECK(hipGetLastError());
        double bb[6];
        IXH_HIP_CHECK(hipMemcpyAsync(bb, h->bbox.ptr, sizeof(bb), hipMemcpyDeviceToHost, stream));
        IXH_HIP_CHECK(hipStreamSynchronize(stream));
        for (int k = 0; k < 3; k++) {
            double extent = bb[3+k]-bb[k];
            int cells = (int) std::min(64.0, std::max(1.0, std::floor(extent/cell)));
            g.n[k] = cells;
            g.origin[k] = bb[k];
            g.scale[k] = (cells > 1 ? cells/(extent*(1+1e-12)) : 0.0);
        }
    }
    return g;
}

template <typename Real>
void launch_triples(ixh_threebody* h, const ixh_buffers* b, const BoxT& box, double cutoff2, int maxCount, hipStream_t stream,
                    int subsetItems = -1) {
    typedef typename Vec4Of<Real>::type R4;
    TripleArgs<Real> t;
    t.sites = h->sites.ptr;
    t.molecule = h->moleculeDev.ptr;
    t.siteParams = h->paramsDev.ptr;
    t.start = h->start.ptr;
    t.count = h->count.ptr;
    const bool items = (h->numSegments > 0);
    t.itemSite = (items ? h->itemSite.ptr : nullptr);
    t.itemFirst = (items ? h->itemFirst.ptr : nullptr);
    t.itemShares = (items ? h->itemShares.ptr : nullptr);
    t.itemShareBase = (items ? h->itemShareBase.ptr : nullptr);
    t.itemIndex = (subsetItems >= 0 ? h->subsetItems.ptr : (items ? h->itemOrder.ptr : (h->lengthOrderValid ? h->lengthOrder.ptr : nullptr)));
    t.nbrIndex = h->nbrIndex.ptr;
    t.numSites = h->numSites;
    t.periodic = h->periodic;
    t.box = box;
    t.boxR[0] = Real(box.ax);
    t.boxR[1] = Real(box.bx);
    t.boxR[2] = Real(box.by);
    t.boxR[3] = Real(box.cx);
    t.boxR[4] = Real(box.cy);
    t.boxR[5] = Real(box.cz);
    t.boxR[6] = Real(h->periodic ? 1.0/box.ax : 0.0);
    t.boxR[7] = Real(h->periodic ? 1.0/box.by : 0.0);
    t.boxR[8] = Real(h->periodic ? 1.0/box.cz : 0.0);
    t.cutoff2 = cutoff2;
    t.cutoff2r = Real(cutoff2);
    t.band = Real(sizeof(Real) == 4 ? 1e-5 : 1e-12);
    t.scaleFactor = Real(h->scaleFactor);
    t.lambda = Real(h->anyAlchemical ? b->lambda : 1.0);
    t.useDamping = (h->dampingMethod == IXH_DAMPING_TANG_TOENNIES);
    t.dampingScale = Real(h->dampingScale);
    t.invDampingRadius = Real(h->centroidDampingRadius > 0 ? 1.0/h->centroidDampingRadius : 0.0);
    t.capacity = std::max(maxCount, 1);
    t.skinLists = (h->numSegments == 0);
    t.contribution = h->contribution.ptr;
    t.exactImage = (h->periodic && 2*h->cutoff >= std::min(box.ax, std::min(box.by, box.cz))) ? 1 : 0;
    t.bcSameImage = (!h->periodic || 4*h->cutoff < 0.999*std::min(box.ax, std::min(box.by, box.cz))) ? 1 : 0;
    t.siteForce = h->siteForce.ptr;
    t.itemEnergy = h->siteEnergy.ptr;
    t.itemCount = h->siteCount.ptr;
    size_t shared = (size_t) t.capacity*(2*sizeof(R4)+3*sizeof(long long)+3*sizeof(int)+sizeof(Real));
    if (shared > 160*1024)
        throw Error(IXH_ERROR_UNSUPPORTED, "ixh_threebody_execute: a site has too many neighbours for one workgroup");
    bool timed = h->timer.start(stream);
    const int workgroups = (subsetItems >= 0 ? subsetItems : (items ? h->numItems : h->numSites));
    if (workgroups > 0) {
        if (items && (t.useDamping || t.invDampingRadius > 0))
            hipLaunchKernelGGL((k_triple_groups<Real, true>), dim3(workgroups), dim3(64*h->tripleWaves), shared, stream, t);
        else if (items)
            hipLaunchKernelGGL((k_triple_groups<Real, false>), dim3(workgroups), dim3(64*h->tripleWaves), shared, stream, t);
        else if (t.capacity <= 64) {
            // several sites per wavefront (PACKED_GROUP lanes each)
            constexpr int SITES = 64/PACKED_GROUP;
            const size_t packed = (size_t) SITES*t.capacity*(2*sizeof(R4)+3*sizeof(long long)+4*sizeof(int)+sizeof(Real));
            const int blocks = (workgroups+SITES-1)/SITES;
            if (t.useDamping || t.invDampingRadius > 0)
                hipLaunchKernelGGL((k_triples_packed<Real, PACKED_GROUP, true>), dim3(blocks), dim3(64), packed, stream, t, workgroups);
            else
                hipLaunchKernelGGL((k_triples_packed<Real, PACKED_GROUP, false>), dim3(blocks), dim3(64), packed, stream, t, workgroups);
        }
        else
            hipLaunchKernelGGL((k_triples<Real>), dim3(workgroups), dim3(64), shared, stream, t);
    }
    IXH_HIP_CHECK(hipGetLastError());
    if (timed)
        h->timer.stop(stream);
}

// The rest of a plain evaluation once its lists hold: the triples, the site forces onto the particles, the energy.
void finish_plain(ixh_threebody* h, const ixh_buffers* b, const BoxT& box, double cutoff2, int flags, hipStream_t stream) {
    const int n = h->numSites;
    if (h->precision == IXH_DOUBLE)
        launch_triples<double>(h, b, box, cutoff2, h->listMax, stream);
    else
        launch_triples<float>(h, b, box, cutoff2, h->listMax, stream);
    const int groups = grid_of((long long) n*GATHER_LANES, GATHER_THREADS);
    h->groupEnergy.reserve(groups);
    h->groupCount.reserve(groups);
    // the packed triple kernel (lists of at most 64 entries) leaves the sites' own forces to the gather
    const bool packed = (h->listMax <= 64);
    hipLaunchKernelGGL(k_gather_site_forces, dim3(groups), dim3(GATHER_THREADS), 0, stream, (const long long*) h->siteForce.ptr,
                       packed ? (const int*) h->start.ptr : nullptr, packed ? (const int*) h->count.ptr : nullptr,
                       (const long long*) h->contribution.ptr, (const int*) h->revStart.ptr, (const int*) h->revEnd.ptr,
                       (const int*) h->revList.ptr, n, (const int*) h->offsetsDev.ptr, (const int*) h->particlesDev.ptr, b->force,
                       b->force_stride, (flags & IXH_INCLUDE_FORCES) != 0 ? 1 : 0, (const double*) h->siteEnergy.ptr,
                       (const int*) h->siteCount.ptr, h->groupEnergy.ptr, h->groupCount.ptr);
    IXH_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_sites, dim3(1), dim3(1024), 0, stream, h->groupEnergy.ptr, h->groupCount.ptr, groups, b->energy,
                       h->tripleCount.ptr, (flags & IXH_INCLUDE_ENERGY) != 0);
    IXH_HIP_CHECK(hipGetLastError());
    IXH_HIP_CHECK(hipEventRecord(h->done, stream));
}

}  // namespace

namespace ixh {

void threebody_set_search_event(ixh_threebody* h, hipEvent_t event) {
    h->searched = event;
}

void threebody_set_launch_shape(ixh_threebody* h, int waves, int chunkPairs) {
    if (waves != 1 && waves != 2 && waves != 4 && waves != 8)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody: the triple kernel takes 1, 2, 4 or 8 wavefronts per work item");
    if (chunkPairs < 1024 || chunkPairs > (1 << 22) || chunkPairs%64 != 0)
        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody: pairs per work item must be a multiple of 64 in [1024, 2^22]");
    h->tripleWaves = waves;
    if (chunkPairs != h->chunkPairs) {
        h->chunkPairs = chunkPairs;
        if (h->numSegments > 0) {
            DeviceGuard guard(h->device);
            IXH_HIP_CHECK(hipDeviceSynchronize());   // no launch in flight reads the items
            build_items(h);
        }
    }
}

void threebody_launch_shape(const ixh_threebody* h, int* waves, int* chunkPairs) {
    *waves = h->tripleWaves;
    *chunkPairs = h->chunkPairs;
}

}  // namespace ixh

extern "C" {

int ixh_threebody_create(int device, const ixh_threebody_desc* d, ixh_threebody** out) {
    return ixh_threebody_create_segmented(device, d, 0, nullptr, out);
}

int ixh_threebody_create_segmented(int device, const ixh_threebody_desc* d, int num_segments, const int* site_offsets,
                                   ixh_threebody** out) {
    return guarded([&] {
        if (out == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody_create: null output pointer");
        check_threebody(d);
        if (num_segments < 0 || (num_segments > 0 && site_offsets == nullptr))
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody_create_segmented: invalid segments");
        if (num_segments > 0) {
            if (d->periodic)
                throw Error(IXH_ERROR_UNSUPPORTED, "ixh_threebody_create_segmented: segments require a nonperiodic force");
            if (site_offsets[0] != 0 || site_offsets[num_segments] != d->num_sites)
                throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody_create_segmented: the segments must cover the sites");
            for (int g = 0; g < num_segments; g++)
                if (site_offsets[g+1] < site_offsets[g])
                    throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody_create_segmented: decreasing segment offsets");
        }
        DeviceGuard guard(device);
        ixh_threebody* h = new ixh_threebody();
        try {
            h->device = device;
            h->numParticles = d->num_particles;
            h->numSites = d->num_sites;
            h->precision = d->precision;
            h->cutoff = d->cutoff;
            h->dampingMethod = d->damping_method;
            h->periodic = d->periodic;
            int n = h->numSites;
            h->offsets.assign(d->site_offsets, d->site_offsets+n+1);
            h->particles.assign(d->site_particles, d->site_particles+h->offsets[n]);
            h->molecule.resize(n);
            h->alchemical.assign(n, 0);
            for (int i = 0; i < n; i++) {
                // ReferenceInteraxThreeBodyIxn.cpp:48-49
                h->molecule[i] = (d->molecule[i] < 0 ? -1-i : d->molecule[i]);
                h->alchemical[i] = (d->alchemical != nullptr && d->alchemical[i] != 0);
            }
            h->anyAlchemical = std::any_of(h->alchemical.begin(), h->alchemical.end(), [](unsigned char c) { return c != 0; });
            h->offsetsDev.upload(h->offsets.data(), h->offsets.size(), nullptr);
            h->particlesDev.upload(h->particles.data(), h->particles.size(), nullptr);
            h->moleculeDev.upload(h->molecule.data(), h->molecule.size(), nullptr);
            h->sites.reserve(3*(size_t) std::max(n, 1));
            h->count.reserve(std::max(n, 1));
            h->siteEnergy.reserve(std::max(n, 1));
            h->siteCount.reserve(std::max(n, 1));
            h->siteForce.reserve(3*(size_t) std::max(n, 1));
            h->tripleCount.reserve(1);
            if (num_segments > 0) {
                h->numSegments = num_segments;
                h->segOffsets.assign(site_offsets, site_offsets+num_segments+1);
                std::vector<int> end(n), start(n+1, 0);
                long long total = 0;
                for (int g = 0; g < num_segments; g++) {
                    h->maxSegment = std::max(h->maxSegment, site_offsets[g+1]-site_offsets[g]);
                    for (int i = site_offsets[g]; i < site_offsets[g+1]; i++) {
                        end[i] = site_offsets[g+1];
                        start[i] = (int) total;
                        total += site_offsets[g+1]-i-1;
                        if (total > INT32_MAX)
                            throw Error(IXH_ERROR_CAPACITY, "ixh_threebody_create_segmented: the neighbour lists exceed 2^31 entries");
                    }
                }
                start[n] = (int) total;
                build_items(h);
                h->segEnergy.reserve(num_segments);
                h->segCount.reserve(num_segments);
                h->segOffsetsDev.upload(h->segOffsets.data(), h->segOffsets.size(), nullptr);
                h->siteEnd.upload(end.data(), std::max<size_t>(end.size(), 1), nullptr);
                h->start.upload(start.data(), start.size(), nullptr);
                h->nbrIndex.reserve((size_t) std::max<long long>(total, 1));
            }
            IXH_HIP_CHECK(hipEventCreateWithFlags(&h->updateDone, hipEventDisableTiming));
            IXH_HIP_CHECK(hipHostMalloc((void**) &h->hostInvalid, sizeof(int), hipHostMallocDefault));
            h->invalid.reserve(1);
            h->rawSites.reserve(3*(size_t) std::max(h->numSites, 1));
            load_threebody_params(h, d);
        }
        catch (...) {
            if (h->updateDone != nullptr)
                (void) hipEventDestroy(h->updateDone);
            if (h->hostInvalid != nullptr)
                (void) hipHostFree(h->hostInvalid);
            delete h;
            throw;
        }
        *out = h;
    });
}

int ixh_threebody_destroy(ixh_threebody* h) {
    return guarded([&] {
        if (h == nullptr)
            return;
        if (h->hostTotals != nullptr)
            (void) hipHostFree(h->hostTotals);
        if (h->hostFlag != nullptr)
            (void) hipHostFree(h->hostFlag);
        DeviceGuard guard(h->device);
        (void) hipDeviceSynchronize();
        if (h->done != nullptr)
            (void) hipEventDestroy(h->done);
        if (h->checked != nullptr)
            (void) hipEventDestroy(h->checked);
        if (h->updateDone != nullptr)
            (void) hipEventDestroy(h->updateDone);
        if (h->hostInvalid != nullptr)
            (void) hipHostFree(h->hostInvalid);
        delete h;
    });
}

int ixh_threebody_set_parameters(ixh_threebody* h, const ixh_threebody_desc* d) {
    return guarded([&] {
        if (h == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody_set_parameters: null handle");
        check_threebody(d);
        if (d->num_sites != h->numSites)
            throw Error(IXH_ERROR_STATE, "updateParametersInContext: The number of sites has changed");
        if ((d->centroid_damping_radius > 0) != (h->centroidDampingRadius > 0))
            throw Error(IXH_ERROR_STATE, "updateParametersInContext: The centroid damping cannot be switched on or off");
        DeviceGuard guard(h->device);
        // the same device work as ixh_threebody_set_parameters_device, from a copy of the host arrays
        const size_t n = (size_t) h->numSites;
        hipStream_t stream = nullptr;
        if (h->done != nullptr)
            IXH_HIP_CHECK(hipStreamWaitEvent(stream, h->done, 0));
        IXH_HIP_CHECK(hipStreamSynchronize(stream));
        if (n > 0) {
            IXH_HIP_CHECK(hipMemcpy(h->rawSites.ptr, d->alpha, sizeof(double)*n, hipMemcpyHostToDevice));
            IXH_HIP_CHECK(hipMemcpy(h->rawSites.ptr+n, d->frequency, sizeof(double)*n, hipMemcpyHostToDevice));
            IXH_HIP_CHECK(hipMemcpy(h->rawSites.ptr+2*n, d->width, sizeof(double)*n, hipMemcpyHostToDevice));
        }
        h->scaleFactor = d->scale_factor;
        h->dampingScale = d->damping_scale;
        h->centroidDampingRadius = d->centroid_damping_radius;
        update_site_params(h, h->rawSites.ptr, h->rawSites.ptr+n, h->rawSites.ptr+2*n, stream);
        finish_site_params(h, stream);
        IXH_HIP_CHECK(hipStreamSynchronize(stream));
    });
}

int ixh_threebody_set_parameters_device(ixh_threebody* h, const ixh_threebody_device_parameters* d) {
    return guarded([&] {
        if (h == nullptr || d == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody_set_parameters_device: null argument");
        if (h->numSites > 0 && (d->alpha == nullptr || d->frequency == nullptr || d->width == nullptr))
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody_set_parameters_device: missing per-site arrays");
        if ((d->centroid_damping_radius > 0) != (h->centroidDampingRadius > 0))
            throw Error(IXH_ERROR_STATE, "updateParametersInContext: The centroid damping cannot be switched on or off");
        DeviceGuard guard(h->device);
        h->scaleFactor = d->scale_factor;
        h->dampingScale = d->damping_scale;
        h->centroidDampingRadius = d->centroid_damping_radius;
        update_site_params(h, d->alpha, d->frequency, d->width, (hipStream_t) d->stream);
    });
}

int ixh_threebody_execute(ixh_threebody* h, const ixh_buffers* b, int flags) {
    return ixh_threebody_execute_segments(h, b, flags, nullptr);
}

int ixh_threebody_execute_segments(ixh_threebody* h, const ixh_buffers* b, int flags, double* segmentEnergies) {
    return ixh_threebody_execute_subset(h, b, flags, segmentEnergies, nullptr, -1);
}

int ixh_threebody_execute_subset(ixh_threebody* h, const ixh_buffers* b, int flags, double* segmentEnergies, const int* segments,
                                 int numSelected) {
    return guarded([&] {
        if (numSelected >= 0 && (h == nullptr || h->numSegments == 0 || (numSelected > 0 && segments == nullptr)))
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody_execute_subset: a subset needs a segmented handle and a segment list");
        if (h == nullptr || b == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody_execute: null argument");
        if (b->num_particles != h->numParticles)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody_execute: the buffers hold a different number of particles");
        int n = h->numSites;
        if (n == 0)
            return;
        DeviceGuard guard(h->device);
        hipStream_t stream = (hipStream_t) b->stream;
        finish_site_params(h, stream);
        if (h->done == nullptr)
            IXH_HIP_CHECK(hipEventCreateWithFlags(&h->done, hipEventDisableTiming));
        else if (stream != h->lastStream)
            IXH_HIP_CHECK(hipStreamWaitEvent(stream, h->done, 0));
        h->lastStream = stream;
        const float4* corr = (const float4*) b->posq_correction;
        BoxT box = {b->box[0], b->box[3], b->box[4], b->box[6], b->box[7], b->box[8]};
        // plain handles whose cell grid comes from the box prepare the sites of a list build in one launch, after the
        // grid
        const bool gridFromBox = (h->numSegments == 0 && (h->periodic || n <= 4096));
        if (h->numSegments > 0) {
            hipLaunchKernelGGL(k_sites, dim3(grid_of(n, 128)), dim3(128), 0, stream, b->posq, corr, b->precision, h->offsetsDev.ptr,
                               h->particlesDev.ptr, n, h->sites.ptr);
            IXH_HIP_CHECK(hipGetLastError());
        }
        double cutoff2 = h->cutoff*h->cutoff;
        if (h->numSegments > 0) {
            // segmented: static list capacity, no host synchronization
            int subsetSites = -1, subsetItemCount = -1;
            if (numSelected >= 0) {
                size_t needed = numSelected;
                for (int k = 0; k < numSelected; k++) {
                    const int g = segments[k];
                    if (g < 0 || g >= h->numSegments)
                        throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody_execute_subset: segment index out of range");
                    needed += h->segOffsets[g+1]-h->segOffsets[g]+h->segItems[g+1]-h->segItems[g];
                }
                int slot = 0;
                int* staging = h->subsetStaging.acquire(needed, slot);
                int count = 0;
                for (int k = 0; k < numSelected; k++)
                    for (int i = h->segOffsets[segments[k]]; i < h->segOffsets[segments[k]+1]; i++)
                        staging[count++] = i;
                int itemCount = 0;
                int* itemList = staging+count;
                for (int k = 0; k < numSelected; k++)
                    for (int w = h->segItems[segments[k]]; w < h->segItems[segments[k]+1]; w++)
                        itemList[itemCount++] = w;
                std::copy(segments, segments+numSelected, itemList+itemCount);
                h->subsetSites.reserve(std::max(count, 1));
                h->subsetItems.reserve(std::max(itemCount, 1));
                h->subsetSegments.reserve(std::max(numSelected, 1));
                if (count > 0)
                    IXH_HIP_CHECK(hipMemcpyAsync(h->subsetSites.ptr, staging, sizeof(int)*count, hipMemcpyHostToDevice, stream));
                if (itemCount > 0)
                    IXH_HIP_CHECK(hipMemcpyAsync(h->subsetItems.ptr, itemList, sizeof(int)*itemCount, hipMemcpyHostToDevice, stream));
                if (numSelected > 0)
                    IXH_HIP_CHECK(hipMemcpyAsync(h->subsetSegments.ptr, itemList+itemCount, sizeof(int)*numSelected,
                                                 hipMemcpyHostToDevice, stream));
                h->subsetStaging.recorded(slot, stream);
                subsetSites = count;
                subsetItemCount = itemCount;
            }
            const int workgroups = (subsetSites >= 0 ? subsetSites : n);
            if (workgroups > 0)
                hipLaunchKernelGGL(k_neighbors_segments, dim3((workgroups+SEGMENT_SEARCH_WAVES-1)/SEGMENT_SEARCH_WAVES),
                                   dim3(64*SEGMENT_SEARCH_WAVES), 0, stream, h->sites.ptr, h->moleculeDev.ptr, h->siteEnd.ptr, cutoff2,
                                   h->start.ptr, h->count.ptr, h->nbrIndex.ptr,
                                   subsetSites >= 0 ? (const int*) h->subsetSites.ptr : (const int*) nullptr, workgroups);
            IXH_HIP_CHECK(hipGetLastError());
            hipLaunchKernelGGL(k_zero_long, dim3(grid_of(3LL*n, 256)), dim3(256), 0, stream, h->siteForce.ptr, 3LL*n);
            IXH_HIP_CHECK(hipGetLastError());
            if (h->searched != nullptr)
                IXH_HIP_CHECK(hipEventRecord(h->searched, stream));
            if (h->precision == IXH_DOUBLE)
                launch_triples<double>(h, b, box, cutoff2, h->maxSegment, stream, subsetItemCount);
            else
                launch_triples<float>(h, b, box, cutoff2, h->maxSegment, stream, subsetItemCount);
            if (flags & IXH_INCLUDE_FORCES) {
                hipLaunchKernelGGL(k_apply_site_forces, dim3(grid_of(n, 128)), dim3(128), 0, stream, h->siteForce.ptr, n,
                                   h->offsetsDev.ptr, h->particlesDev.ptr, b->force, b->force_stride);
                IXH_HIP_CHECK(hipGetLastError());
            }
            // energies and triple counts per segment, then (whole evaluations) the total over the segments in order
            const int reduced = (subsetSites >= 0 ? numSelected : h->numSegments);
            if (reduced > 0) {
                hipLaunchKernelGGL(k_reduce_site_segments, dim3(reduced), dim3(256), 0, stream, (const double*) h->siteEnergy.ptr,
                                   (const int*) h->siteCount.ptr, (const int*) h->segSharesDev.ptr,
                                   subsetSites >= 0 ? (const int*) h->subsetSegments.ptr : (const int*) nullptr,
                                   h->segEnergy.ptr, h->segCount.ptr, segmentEnergies);
                IXH_HIP_CHECK(hipGetLastError());
            }
            if (subsetSites < 0) {
                hipLaunchKernelGGL(k_reduce_sites, dim3(1), dim3(1024), 0, stream, h->segEnergy.ptr, h->segCount.ptr, h->numSegments,
                                   b->energy, h->tripleCount.ptr, (flags & IXH_INCLUDE_ENERGY) != 0);
                IXH_HIP_CHECK(hipGetLastError());
            }
            IXH_HIP_CHECK(hipEventRecord(h->done, stream));
            return;
        }
        // Plain handles: every call computes the centroids and checks the lists (k_sites_check); the lists (within cutoff
        // + skin) are rebuilt only when a site has moved more than half the skin since they were built.  With
        // IXH_DEFER_LIST_CHECK the call does not wait for the check: ixh_threebody_list_was_valid tells afterwards.
        const double listCutoff = h->cutoff+THREE_BODY_SKIN;
        if (h->hostFlag == nullptr) {
            IXH_HIP_CHECK(hipHostMalloc((void**) &h->hostFlag, 2*sizeof(int), hipHostMallocCoherent));
            h->hostFlag[0] = h->hostFlag[1] = 0;
        }
        if (h->checked == nullptr)
            IXH_HIP_CHECK(hipEventCreateWithFlags(&h->checked, hipEventDisableTiming));
        const bool listed = (h->listBuilt && h->listSites == n);
        const bool defer = listed && (flags & IXH_DEFER_LIST_CHECK) != 0;
        const int f = h->frame ^ 1;
        h->frame = f;
        h->deferredCheck = false;
        hipExtLaunchKernelGGL(k_sites_check, dim3(grid_of(3*n, 256)), dim3(256), 0, stream, nullptr, defer ? h->checked : nullptr, 0,
                              b->posq, corr, b->precision, (const int*) h->offsetsDev.ptr, (const int*) h->particlesDev.ptr, n,
                              h->sites.ptr, listed ? (const double*) h->buildSites.ptr : (const double*) nullptr, h->periodic, box,
                              0.25*THREE_BODY_SKIN*THREE_BODY_SKIN, h->hostFlag+f, h->hostFlag+(f ^ 1), h->siteForce.ptr);
        IXH_HIP_CHECK(hipGetLastError());
        bool rebuild = !listed;
        if (defer)
            h->deferredCheck = true;
        else if (listed) {
            IXH_HIP_CHECK(hipStreamSynchronize(stream));
            rebuild = (h->hostFlag[f] != 0);
        }
        if (!rebuild) {
            finish_plain(h, b, box, cutoff2, flags, stream);
            return;
        }
        CellGrid g = make_grid(h, b->box, listCutoff, stream);
        const int numCells = g.n[0]*g.n[1]*g.n[2];
        h->cellKeys.reserve(n);
        h->cellKeysSorted.reserve(n);
        h->cellValues.reserve(n);
        h->sortedSites.reserve(n);
        h->cellStart.reserve(numCells);
        h->cellEnd.reserve(numCells);
        // the sites in cell order: one cell needs no sort; grids of a few sites per cell a counting sort (three small
        // launches); other grids the radix sort
        const bool identity = (gridFromBox && numCells == 1);
        const bool counting = (gridFromBox && !identity && numCells <= 65536 && (long long) n <= 64LL*numCells);
        if (counting) {
            const size_t had = h->cellCount.count;
            h->cellCount.reserve(numCells);
            h->cellCursor.reserve(numCells);
            if (h->cellCount.count != had || had == 0)
                hipLaunchKernelGGL(k_fill_int, dim3(grid_of((int) h->cellCount.count, 256)), dim3(256), 0, stream, h->cellCount.ptr,
                                   (int) h->cellCount.count, 0);
        }
        if (gridFromBox)
            hipLaunchKernelGGL(k_prepare_sites, dim3(grid_of(std::max(3*n, numCells), 256)), dim3(256), 0, stream, b->posq, corr,
                               b->precision, (const int*) h->offsetsDev.ptr, (const int*) h->particlesDev.ptr, n, h->sites.ptr, g,
                               h->cellKeys.ptr, h->cellValues.ptr, h->cellStart.ptr, h->cellEnd.ptr, numCells, h->siteForce.ptr,
                               counting ? h->cellCount.ptr : (int*) nullptr, identity ? h->sortedSites.ptr : (int*) nullptr);
        else
            hipLaunchKernelGGL(k_site_cells, dim3(grid_of(n, 256)), dim3(256), 0, stream, h->sites.ptr, n, g, h->cellKeys.ptr,
                               h->cellValues.ptr);
        IXH_HIP_CHECK(hipGetLastError());
        int bits = 1;
        while ((1LL << bits) < numCells)
            bits++;
        // every buffer of the search is sized before the first launch (a reallocation frees memory in use)
        h->count.reserve(n);
        h->listTotals.reserve(2);
        if (h->hostTotals == nullptr)
            IXH_HIP_CHECK(hipHostMalloc((void**) &h->hostTotals, 2*sizeof(int), hipHostMallocDefault));
        size_t tempBytes = 0, maxBytes = 0;
        IXH_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tempBytes, h->cellKeys.ptr, h->cellKeysSorted.ptr, h->cellValues.ptr,
                                                h->sortedSites.ptr, (size_t) n, 0, bits, stream));
        IXH_HIP_CHECK(rocprim::reduce(nullptr, maxBytes, h->count.ptr, h->listTotals.ptr, 0, (size_t) n, rocprim::maximum<int>(),
                                      stream));
        h->sortTemp.reserve(std::max(tempBytes, maxBytes));
        if (counting) {
            hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, stream, h->cellCount.ptr, numCells, h->cellStart.ptr, h->cellEnd.ptr,
                               h->cellCursor.ptr);
            hipLaunchKernelGGL(k_scatter_sites, dim3(grid_of(n, 256)), dim3(256), 0, stream, (const unsigned*) h->cellKeys.ptr, n,
                               h->cellCursor.ptr, h->sortedSites.ptr);
            hipLaunchKernelGGL(k_sort_cells, dim3(grid_of(numCells, 64)), dim3(64), 0, stream, (const int*) h->cellStart.ptr,
                               (const int*) h->cellEnd.ptr, numCells, h->sortedSites.ptr);
        }
        else if (!identity) {
            IXH_HIP_CHECK(rocprim::radix_sort_pairs(h->sortTemp.ptr, tempBytes, h->cellKeys.ptr, h->cellKeysSorted.ptr,
                                                    h->cellValues.ptr, h->sortedSites.ptr, (size_t) n, 0, bits, stream));
            if (!gridFromBox) {
                hipLaunchKernelGGL(k_fill_int, dim3(grid_of(numCells, 256)), dim3(256), 0, stream, h->cellStart.ptr, numCells, 0);
                hipLaunchKernelGGL(k_fill_int, dim3(grid_of(numCells, 256)), dim3(256), 0, stream, h->cellEnd.ptr, numCells, 0);
            }
            hipLaunchKernelGGL(k_cell_ranges, dim3(grid_of(n, 256)), dim3(256), 0, stream, h->cellKeysSorted.ptr, n, h->cellStart.ptr,
                               h->cellEnd.ptr);
        }
        IXH_HIP_CHECK(hipGetLastError());
        const int exactImage = (h->periodic && 2*listCutoff >= std::min(box.ax, std::min(box.by, box.cz))) ? 1 : 0;
        // Lists of a fixed capacity per site, written in one pass; only the longest list comes back to the host (it
        // sizes the triple kernel's LDS), and a list longer than the capacity makes the search run again with more.
        int maxCount = 0;
        for (int attempt = 0; attempt < 2; attempt++) {
            if (h->listCapacity <= 0 || h->listSites != n) {
                h->listCapacity = std::max(h->listCapacity, 64);
                h->listSites = n;
                h->start.reserve(n);
                h->nbrIndex.reserve((size_t) n*h->listCapacity);
                hipLaunchKernelGGL(k_strided_starts, dim3(grid_of(n, 256)), dim3(256), 0, stream, h->start.ptr, n, h->listCapacity);
                IXH_HIP_CHECK(hipGetLastError());
            }
            hipLaunchKernelGGL(k_neighbors, dim3(n), dim3(64), 0, stream, h->sites.ptr, h->moleculeDev.ptr, n, h->periodic, box,
                               listCutoff*listCutoff, g, exactImage, h->cellStart.ptr, h->cellEnd.ptr, h->sortedSites.ptr, h->count.ptr,
                               h->listCapacity, h->nbrIndex.ptr);
            IXH_HIP_CHECK(hipGetLastError());
            IXH_HIP_CHECK(rocprim::reduce(h->sortTemp.ptr, maxBytes, h->count.ptr, h->listTotals.ptr, 0, (size_t) n,
                                          rocprim::maximum<int>(), stream));
            IXH_HIP_CHECK(hipMemcpyAsync(h->hostTotals, h->listTotals.ptr, sizeof(int), hipMemcpyDeviceToHost, stream));
            IXH_HIP_CHECK(hipStreamSynchronize(stream));
            maxCount = h->hostTotals[0];
            if (maxCount <= h->listCapacity)
                break;
            int capacity = 64;
            while (capacity < maxCount+maxCount/4)
                capacity *= 2;
            h->listCapacity = capacity;
            h->listSites = -1;   // reallocate and repeat the search
        }
        if (maxCount > h->listCapacity)
            throw Error(IXH_ERROR_HIP, "ixh_threebody_execute: neighbour list capacity not reached after a resize");
        // the positions the lists hold for, for the checks of the next calls
        h->buildSites.reserve(3*(size_t) n);
        IXH_HIP_CHECK(hipMemcpyAsync(h->buildSites.ptr, h->sites.ptr, sizeof(double)*3*(size_t) n, hipMemcpyDeviceToDevice, stream));
        h->listBuilt = true;
        h->listMax = maxCount;
        h->lengthOrderValid = false;
        if (maxCount <= 64) {
            // the packed triple kernel's site order: longest list first
            h->lengthOrder.reserve(n);
            h->lengthKeys.reserve(n);
            h->siteIota.reserve(n);
            hipLaunchKernelGGL(k_iota_int, dim3(grid_of(n, 256)), dim3(256), 0, stream, h->siteIota.ptr, n);
            size_t orderBytes = 0;
            IXH_HIP_CHECK(rocprim::radix_sort_pairs_desc(nullptr, orderBytes, h->count.ptr, h->lengthKeys.ptr, h->siteIota.ptr,
                                                         h->lengthOrder.ptr, (size_t) n, 0, 7, stream));
            if (orderBytes > h->sortTemp.count) {
                IXH_HIP_CHECK(hipStreamSynchronize(stream));   // the old buffer may still be in use by the search
                h->sortTemp.reserve(orderBytes);
            }
            IXH_HIP_CHECK(rocprim::radix_sort_pairs_desc(h->sortTemp.ptr, orderBytes, h->count.ptr, h->lengthKeys.ptr, h->siteIota.ptr,
                                                         h->lengthOrder.ptr, (size_t) n, 0, 7, stream));
            h->lengthOrderValid = true;
        }
        // the reverse lists: which list entries name each site
        {
            const long long entries = (long long) n*h->listCapacity;
            h->contribution.reserve(3*(size_t) entries);
            h->revList.reserve((size_t) entries);
            h->revStart.reserve(n);
            h->revEnd.reserve(n);
            h->revCursor.reserve(n);
            if (h->revCount.count < (size_t) n || h->revCount.ptr == nullptr) {
                h->revCount.reserve(n);
                hipLaunchKernelGGL(k_fill_int, dim3(grid_of(n, 256)), dim3(256), 0, stream, h->revCount.ptr, n, 0);
            }
            const int grid = (int) std::max(1LL, (entries+255)/256);
            hipLaunchKernelGGL(k_reverse_lists, dim3(grid), dim3(256), 0, stream, (const int*) h->nbrIndex.ptr, (const int*) h->count.ptr, n,
                               h->listCapacity, h->revCount.ptr, (int*) nullptr, (int*) nullptr);
            hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, stream, h->revCount.ptr, n, h->revStart.ptr, h->revEnd.ptr,
                               h->revCursor.ptr);
            hipLaunchKernelGGL(k_reverse_lists, dim3(grid), dim3(256), 0, stream, (const int*) h->nbrIndex.ptr, (const int*) h->count.ptr, n,
                               h->listCapacity, h->revCount.ptr, h->revCursor.ptr, h->revList.ptr);
            IXH_HIP_CHECK(hipGetLastError());
        }
        finish_plain(h, b, box, cutoff2, flags, stream);
    });
}

int ixh_threebody_kernel_time(ixh_threebody* h, double* milliseconds, long long* launches) {
    return guarded([&] {
        if (h == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody_kernel_time: null handle");
        DeviceGuard guard(h->device);
        h->timer.collect();
        if (milliseconds != nullptr)
            *milliseconds = h->timer.accumulated;
        if (launches != nullptr)
            *launches = h->timer.launches;
    });
}

int ixh_threebody_reset_kernel_time(ixh_threebody* h) {
    return guarded([&] {
        if (h == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody_reset_kernel_time: null handle");
        DeviceGuard guard(h->device);
        h->timer.reset();
    });
}

int ixh_threebody_list_was_valid(ixh_threebody* h, int* valid) {
    return guarded([&] {
        if (h == nullptr || valid == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody_list_was_valid: null argument");
        *valid = 1;
        if (h->deferredCheck) {
            DeviceGuard guard(h->device);
            IXH_HIP_CHECK(hipEventSynchronize(h->checked));
            *valid = (h->hostFlag[h->frame] == 0);
        }
    });
}

int ixh_threebody_last_triple_count(ixh_threebody* h, long long* count) {
    return guarded([&] {
        if (h == nullptr || count == nullptr)
            throw Error(IXH_ERROR_INVALID_ARGUMENT, "ixh_threebody_last_triple_count: null argument");
        DeviceGuard guard(h->device);
        IXH_HIP_CHECK(hipStreamSynchronize(h->lastStream));
        IXH_HIP_CHECK(hipMemcpy(count, h->tripleCount.ptr, sizeof(long long), hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
