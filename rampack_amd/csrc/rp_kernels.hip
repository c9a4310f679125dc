// rp_kernels.hip -- hand-written gfx950 kernels for the rampack MC sweep.
//
//   * k_pair_decisions   batched FP64 pair-overlap decisions (the decision function of
//                        SpherocylinderTraits::overlapBetween / SphereTraits, SpherocylinderTraits.cpp:28-46,
//                        SphereTraits.cpp:56-65) on explicit (pose, pose, translation) records
//   * k_bin / k_scan / k_scatter / k_gather
//                        counting-sort cell-list rebuild over coalesced SoA records, replacing the linked-list
//                        NeighbourGrid insertion (Packing.cpp:858-929, NeighbourGrid.cpp:219-227); the order
//                        inside a cell is the stable order of the source arrays, so rebuilds are deterministic
//   * k_count_overlaps   half-shell overlap count with early exit (Packing.cpp:431-559), one workgroup per cell
//                        with the cell and its 13 positive neighbours staged in LDS
//   * k_sweep            checkerboard MC sweep: one workgroup (1, 2 or 4 wavefronts) per active cell, the 27-cell
//                        stencil staged in LDS, early-exit overlap voting after every 64*waves candidates, a Philox
//                        stream per move, moves that leave the cell rejected (the ActiveDomain rule of
//                        Packing.cpp:147-148)
#include "rp_kernels.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cfloat>
#include <mutex>

namespace rp {

namespace {

constexpr int WAVE = 64;

__device__ __forceinline__ V3 load3(const double *x, const double *y, const double *z, int i) {
    return v3(x[i], y[i], z[i]);
}

// Cell coordinates and wrap bits of `pos` in grid `g`.
//   exact: NeighbourGrid::positionToCellNo (NeighbourGrid.cpp:13-34), relative coordinate clamped into [0, 1)
//   shifted: y = x - shift + w with w = (x < shift), so y is in [0, 1); a particle with w = 1 always sits in
//            the last cell, which keeps every image offset between stencil neighbours in {-1, 0, 1}.
__device__ __forceinline__ void bin_position(const Grid &g, const V3 &pos, int c[3], unsigned &wbits) {
    V3 rel = mulv(g.inv, pos);
    double r[3] = {rel.x, rel.y, rel.z};
    wbits = 0;
    for (int d = 0; d < 3; d++) {
        double x = r[d];
        int n = g.n[d];
        int cc;
        if (g.exact) {
            const double EPS = DBL_EPSILON * 10;
            if (x < 0) x = 0;
            else if (x >= 1) x = 1 - EPS;
            cc = (int)(x / g.relCell[d]);
        } else {
            int w = x < g.shift[d] ? 1 : 0;
            double y = (x - g.shift[d]) + (w ? 1.0 : 0.0);
            cc = (int)(y * (double)n);
            if (w) cc = n - 1;
            wbits |= (unsigned)w << d;
        }
        if (cc < 0) cc = 0;
        if (cc > n - 1) cc = n - 1;
        c[d] = cc;
    }
}

__device__ __forceinline__ int cell_index(const Grid &g, int cx, int cy, int cz) {
    return cx + g.n[0] * (cy + g.n[1] * cz);
}

// Neighbour cell coordinate along one axis with periodic wrap; e receives the image shift (-1, 0, +1).
__device__ __forceinline__ int wrap_axis(int c, int o, int n, int &e) {
    int cc = c + o;
    e = 0;
    if (cc < 0) { cc += n; e = -1; }
    else if (cc >= n) { cc -= n; e = 1; }
    return cc;
}

template <int KIND>
__device__ __forceinline__ bool pair_overlap(const PairConsts &k, const SCFirst &f, const V3 &p2, const V3 &a2,
                                             const V3 &T) {
    if (KIND == 0) return sphere_overlap(k, f.pos, p2, T);
    return sc_overlap(k, f, p2, a2, T);
}

template <int KIND>
__device__ __forceinline__ SCFirst make_first(const PairConsts &k, const V3 &pos, const V3 &axis) {
    if (KIND != 1) {
        SCFirst f;
        f.pos = pos;
        return f;
    }
    return sc_first(k, pos, axis);
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------------
// Pair decisions on records pos1[3] O1[9] pos2[3] O2[9] T[3] (row-major orientations)

__global__ __launch_bounds__(256) void k_pair_decisions(PairConsts k, size_t n, const double *__restrict__ rec,
                                                        uint8_t *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *q = rec + 27 * i;
    V3 p1 = v3(q[0], q[1], q[2]);
    V3 a1 = v3(q[3 + 2], q[3 + 5], q[3 + 8]);
    V3 p2 = v3(q[12], q[13], q[14]);
    V3 a2 = v3(q[15 + 2], q[15 + 5], q[15 + 8]);
    V3 T = v3(q[24], q[25], q[26]);
    bool o = false;
    if (k.kind == 0) {
        o = sphere_overlap(k, p1, p2, T);
    } else if (k.kind == 1) {
        SCFirst f = sc_first(k, p1, a1);
        o = sc_overlap(k, f, p2, a2, T);
    } else {
        // bead centres pos + O * c (PolysphereTraits::SphereData::centreForShape, PolysphereTraits.cpp:93-95);
        // bead pairs as PolysphereTraits::HardInteraction::overlapBetween (PolysphereTraits.cpp:97-106)
        for (int b = 0; b < k.nc && !o; b++) {
            const double *cb = k.beads + 4 * b;
            const V3 c2 = add(p2, mulv(q + 15, v3(cb[0], cb[1], cb[2])));
            for (int a = 0; a < k.nc && !o; a++) {
                const double *ca = k.beads + 4 * a;
                const V3 c1 = add(p1, mulv(q + 3, v3(ca[0], ca[1], ca[2])));
                o = bead_overlap(ca[3], cb[3], c1, c2, T);
            }
        }
    }
    out[i] = o ? 1 : 0;
}

// Absolute bead centres of every particle (Packing::recalculateAbsoluteInteractionCentres, Packing.cpp:1036-1047):
// item i * nc + c is position + orientation * centre_c, wrapped by the periodic correction of that sum; its tag is
// the item index (the reference's centre index, the order in which its grid lists centres).
__global__ __launch_bounds__(256) void k_centres(Particles src, int n, PairConsts k, Grid g, Particles dst) {
    const int item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= n * k.nc) return;
    const int i = item / k.nc, c = item - i * k.nc;
    double O[9];
    for (int e = 0; e < 9; e++) O[e] = src.o[e][i];
    const double *cb = k.beads + 4 * c;
    const V3 pos = translate_wrapped(g.dims, g.inv, load3(src.x, src.y, src.z, i), mulv(O, v3(cb[0], cb[1], cb[2])));
    dst.x[item] = pos.x;
    dst.y[item] = pos.y;
    dst.z[item] = pos.z;
    dst.tag[item] = item;
}

hipError_t launch_centres(hipStream_t s, const Particles &src, int n, const PairConsts &k, const Grid &g,
                          const Particles &dst) {
    const size_t items = (size_t)n * k.nc;
    if (items == 0) return hipSuccess;
    hipLaunchKernelGGL(k_centres, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, s, src, n, k, g, dst);
    return hipGetLastError();
}

hipError_t launch_pair_decisions(hipStream_t s, PairConsts k, size_t n, const double *rec, uint8_t *out) {
    if (n == 0) return hipSuccess;
    unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_pair_decisions, dim3(blocks), dim3(256), 0, s, k, n, rec, out);
    return hipGetLastError();
}

// Guarded FP32 decisions of the same records.  Coordinates are taken relative to r0 = floor(pos1) (FP32 roundings
// of FP64 differences, as the staged values of a sweep would be); a pair the FP32 bounds decide keeps that decision,
// any other pair takes the FP64 decision.  Per record: out = the decision; code = GuardResult of the deciding FP32
// stage | 4 if the segment stage ran | 8 if the decision differs from the FP64 one (must never happen); errC =
// |d2f - d2_64|; errS = |segment d2 in FP32 - in FP64| at the FP32 (s, t), or -1 without a segment stage.
// VERIFY: also evaluate every pair in FP64 for the codes and error statistics; otherwise FP64 runs only for the
// pairs the FP32 stages leave undecided.
template <bool VERIFY>
__global__ __launch_bounds__(256) void k_pair_decisions_guarded(PairConsts k, size_t n, const double *__restrict__ rec,
                                                                uint8_t *__restrict__ out, uint8_t *__restrict__ code,
                                                                float *__restrict__ errC, float *__restrict__ errS) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *q = rec + 27 * i;
    const V3 p1 = v3(q[0], q[1], q[2]);
    const V3 a1 = v3(q[3 + 2], q[3 + 5], q[3 + 8]);
    const V3 p2 = v3(q[12], q[13], q[14]);
    const V3 a2 = v3(q[15 + 2], q[15 + 5], q[15 + 8]);
    const V3 T = v3(q[24], q[25], q[26]);
    const V3 p2bc = add(p2, T);
    // FP64 reference decision (and values, for the statistics)
    auto fp64 = [&](double &d2_64, double &seg64) {
        const V3 dd = sub(p2bc, p1);
        d2_64 = dot(dd, dd);
        seg64 = -1;
        if (k.kind == 0) return sphere_overlap(k, p1, p2, T);
        const SCFirst f1 = sc_first(k, p1, a1);
        if (VERIFY && !(d2_64 < k.contact2) && d2_64 < k.range2)
            seg64 = segment_distance2(f1.capA, f1.u, f1.a, add(p2bc, scal(a2, -k.halfLen)),
                                      add(p2bc, scal(a2, k.halfLen)));
        return sc_overlap(k, f1, p2, a2, T);
    };
    double d2_64 = 0, seg64 = -1;
    bool o64 = false;
    if (VERIFY) o64 = fp64(d2_64, seg64);
    // guarded FP32
    const V3 r0 = v3(floor(p1.x), floor(p1.y), floor(p1.z));
    const F3 fm = to_f3(sub(p1, r0)), fj = to_f3(sub(p2bc, r0));
    const F3 em = round_err(fm), ej = round_err(fj);
    float bound;
    const float d2f = guard_centre(fm, em, fj, ej, bound);
    const double lo = (double)d2f - (double)bound, hi = (double)d2f + (double)bound;
    int g = GUARD_UNDECIDED, flags = 0;
    float segf = -1;
    if (hi < k.contact2) g = GUARD_OVERLAP;
    else if (k.kind == 0 && lo >= k.contact2) g = GUARD_NO;
    else if (k.kind == 1 && lo >= k.range2) g = GUARD_NO;
    else if (k.kind == 1 && lo >= k.contact2 && hi < k.range2) {
        flags |= 4;
        const float h = (float)k.halfLen;
        const float eh = (float)fabs((double)h - k.halfLen);
        g = guard_segment(k.contact2, fm, to_f3(a1), em, fj, to_f3(a2), ej, h, eh, segf);
    }
    if (!VERIFY) {
        if (g == GUARD_UNDECIDED) o64 = fp64(d2_64, seg64);
        out[i] = (g == GUARD_UNDECIDED ? o64 : g == GUARD_OVERLAP) ? 1 : 0;
        return;
    }
    const bool o = g == GUARD_UNDECIDED ? o64 : g == GUARD_OVERLAP;
    if (o != o64) flags |= 8;
    out[i] = o ? 1 : 0;
    code[i] = (uint8_t)(g | flags);
    errC[i] = (float)fabs((double)d2f - d2_64);
    errS[i] = (flags & 4) && seg64 >= 0 && segf >= 0 ? (float)fabs((double)segf - seg64) : -1.f;
}

hipError_t launch_pair_decisions_guarded(hipStream_t s, PairConsts k, size_t n, const double *rec, uint8_t *out,
                                         uint8_t *code, float *errC, float *errS) {
    if (n == 0) return hipSuccess;
    unsigned blocks = (unsigned)((n + 255) / 256);
    if (code)
        hipLaunchKernelGGL(k_pair_decisions_guarded<true>, dim3(blocks), dim3(256), 0, s, k, n, rec, out, code, errC,
                           errS);
    else
        hipLaunchKernelGGL(k_pair_decisions_guarded<false>, dim3(blocks), dim3(256), 0, s, k, n, rec, out, code,
                           errC, errS);
    return hipGetLastError();
}

// Statistics of n guarded decisions (codes and errors of k_pair_decisions_guarded<true>) into counters: [0] guard-band
// fallbacks, [1] segment stages, [2] segment-stage fallbacks, [3] mismatches, [4] largest centre error and [5] largest
// segment error as float bits (non-negative floats order as their bits).  The counters start at zero.
constexpr int GUARD_STATS_THREADS = 256;

__global__ __launch_bounds__(GUARD_STATS_THREADS) void k_guard_stats(size_t n, const uint8_t *__restrict__ code,
                                                                     const float *__restrict__ errC,
                                                                     const float *__restrict__ errS,
                                                                     unsigned long long *counters) {
    unsigned long long c[4] = {0, 0, 0, 0};
    float mc = 0.f, ms = 0.f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint8_t v = code[i];
        c[0] += (v & 3) == GUARD_UNDECIDED ? 1 : 0;
        c[1] += (v & 4) ? 1 : 0;
        c[2] += (v & 4) && (v & 3) == GUARD_UNDECIDED ? 1 : 0;
        c[3] += (v & 8) ? 1 : 0;
        mc = fmaxf(mc, errC[i]);
        if (errS[i] >= 0.f) ms = fmaxf(ms, errS[i]);
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        for (int q = 0; q < 4; q++) c[q] += __shfl_xor(c[q], off);
        mc = fmaxf(mc, __shfl_xor(mc, off));
        ms = fmaxf(ms, __shfl_xor(ms, off));
    }
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        for (int q = 0; q < 4; q++)
            if (c[q]) atomicAdd(&counters[q], c[q]);
        atomicMax(&counters[4], (unsigned long long)__float_as_uint(mc));
        atomicMax(&counters[5], (unsigned long long)__float_as_uint(ms));
    }
}

hipError_t launch_guard_stats(hipStream_t s, size_t n, const uint8_t *code, const float *errC, const float *errS,
                              unsigned long long *counters) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<size_t>((n + GUARD_STATS_THREADS - 1) / GUARD_STATS_THREADS, 1024);
    hipLaunchKernelGGL(k_guard_stats, dim3(blocks), dim3(GUARD_STATS_THREADS), 0, s, n, code, errC, errS, counters);
    return hipGetLastError();
}

// Fixed device memory of the allocation-free guarded pair decisions: 256 bytes of counters, then the centre and
// segment errors (GUARD_SCRATCH_PAIRS floats each) and the codes (GUARD_SCRATCH_PAIRS bytes).  Every GPU the library
// runs on holds its own copy for the life of the process; its device address comes from hipGetSymbolAddress, once
// per device (the array's name in host code is not a device address).
static __device__ __align__(256) unsigned char scratch[GUARD_SCRATCH_BYTES];

namespace {
constexpr int SCRATCH_DEVICES = 64;
std::mutex scratchMutex;
void *scratchPtr[SCRATCH_DEVICES] = {};
}  // namespace

hipError_t guard_scratch(GuardScratch &out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= SCRATCH_DEVICES) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lock(scratchMutex);
    if (!scratchPtr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidValue;
        scratchPtr[dev] = p;
    }
    unsigned char *base = static_cast<unsigned char *>(scratchPtr[dev]);
    out.counters = reinterpret_cast<unsigned long long *>(base);
    out.errC = reinterpret_cast<float *>(base + 256);
    out.errS = out.errC + GUARD_SCRATCH_PAIRS;
    out.code = reinterpret_cast<uint8_t *>(out.errS + GUARD_SCRATCH_PAIRS);
    return hipSuccess;
}

__global__ __launch_bounds__(256) void k_pair_energies(PairConsts k, size_t n, const double *__restrict__ rec,
                                                       double *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *q = rec + 27 * i;
    V3 p1 = v3(q[0], q[1], q[2]);
    V3 a1 = v3(q[5], q[8], q[11]);
    V3 p2 = v3(q[12], q[13], q[14]);
    V3 a2 = v3(q[17], q[20], q[23]);
    V3 T = v3(q[24], q[25], q[26]);
    SCFirst f;
    if (k.kind == 0) f.pos = p1;
    else f = sc_first(k, p1, a1);
    out[i] = soft_energy(k, f, p2, a2, T);
}

hipError_t launch_pair_energies(hipStream_t s, PairConsts k, size_t n, const double *rec, double *out) {
    if (n == 0) return hipSuccess;
    unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_pair_energies, dim3(blocks), dim3(256), 0, s, k, n, rec, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------------------
// Fill kernels (used instead of hipMemsetAsync so the cycle stays capturable in a graph)

__global__ void k_fill_u64(unsigned long long *p, size_t n, unsigned long long v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i32(int *p, size_t n, int v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_u64(hipStream_t s, unsigned long long *p, size_t n, unsigned long long v) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// Clears the overlap count and the stage-overflow counter before an overlap count (one launch for both).
__global__ void k_clear_count(unsigned long long *counters) {
    if (threadIdx.x == 0) counters[CNT_OVERLAPS] = 0;
    if (threadIdx.x == 1) counters[CNT_STAGE_OVERFLOW] = 0;
}

hipError_t launch_clear_count(hipStream_t s, unsigned long long *counters) {
    hipLaunchKernelGGL(k_clear_count, dim3(1), dim3(64), 0, s, counters);
    return hipGetLastError();
}

hipError_t launch_fill_i32(hipStream_t s, int *p, size_t n, int v) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------------------
// Cell-list rebuild (counting sort)

__global__ __launch_bounds__(256) void k_bin(Grid g, int n, Particles src, int rescale, M3 oldInv,
                                             double *__restrict__ tmpPos, int *__restrict__ tmpCell,
                                             int *__restrict__ tmpSlot, uint8_t *__restrict__ tmpW) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    int cell = -1;
    if (valid) {
        V3 pos = load3(src.x, src.y, src.z, i);
        if (rescale) pos = mulv(g.dims, mulv(oldInv.m, pos));
        tmpPos[i] = pos.x;
        tmpPos[n + i] = pos.y;
        tmpPos[2 * n + i] = pos.z;
        int c[3];
        unsigned w;
        bin_position(g, pos, c, w);
        cell = cell_index(g, c[0], c[1], c[2]);
        tmpCell[i] = cell;
        tmpW[i] = (uint8_t)w;
    }
    // The particles arrive in the previous grid's cell order, so a wavefront's lanes fall into a few cells: one
    // counter update per distinct cell, each lane's slot its rank among the lanes of that cell.  (Slots order the
    // members only until the gather, which ranks them by source index.)
    unsigned long long active = __ballot(valid);
    int slot = 0;
    while (active) {
        const int leader = __ffsll((long long)active) - 1;
        const int lc = __builtin_amdgcn_readlane(cell, leader);
        const unsigned long long same = __ballot(valid && cell == lc);
        int base = 0;
        if ((int)(threadIdx.x & (WAVE - 1)) == leader) base = atomicAdd(&g.cellCount[lc], (int)__popcll(same));
        base = __builtin_amdgcn_readlane(base, leader);
        if (valid && cell == lc)
            slot = base + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(same >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((unsigned)same, 0u));
        active &= ~same;
    }
    if (valid) tmpSlot[i] = slot;
}

// Exclusive scan of the cell counts into cellStart (and cellStart[ncell] = n); the counts are cleared for the next
// rebuild.  Up to SCAN_ONE_BLOCK cells: one workgroup, a contiguous chunk per thread; larger grids (bead grids,
// million-particle packings): SCAN_TILE cells per workgroup -- tile sums, a scan of the tile sums, then each tile
// scanned with its offset.  Within a workgroup the thread sums are scanned by wavefront shuffles and one pass over
// the wavefront totals.
constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_ONE_BLOCK = 16384;
constexpr int SCAN_ITEMS = 8;                            // cells per thread in a tile
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;      // cells per workgroup
constexpr int SCAN_MAX_TILES = SCAN_THREADS;              // tile sums scanned by one workgroup

__device__ __forceinline__ int wave_inclusive_scan(int v) {
    const int lane = threadIdx.x & (WAVE - 1);
    for (int off = 1; off < WAVE; off <<= 1) {
        const int t = __shfl_up(v, off);
        if (lane >= off) v += t;
    }
    return v;
}

// Exclusive prefix of v over the workgroup's threads (SCAN_THREADS); *total receives the sum.
__device__ __forceinline__ int block_exclusive_scan(int v, int *waveTot, int *total) {
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    constexpr int NW = SCAN_THREADS / WAVE;
    const int incl = wave_inclusive_scan(v);
    if (lane == WAVE - 1) waveTot[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        int w = lane < NW ? waveTot[lane] : 0;
        w = wave_inclusive_scan(w);
        if (lane < NW) waveTot[lane] = w;
    }
    __syncthreads();
    *total = waveTot[NW - 1];
    return incl - v + (wave > 0 ? waveTot[wave - 1] : 0);
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan(int ncell, int *__restrict__ count, int *__restrict__ start) {
    __shared__ int waveTot[SCAN_THREADS / WAVE];
    const int t = threadIdx.x;
    const int per = (ncell + SCAN_THREADS - 1) / SCAN_THREADS;
    const int b = min(t * per, ncell), e = min(b + per, ncell);
    int sum = 0;
    for (int i = b; i < e; i++) sum += count[i];
    int total = 0;
    int run = block_exclusive_scan(sum, waveTot, &total);
    for (int i = b; i < e; i++) {
        const int c = count[i];
        start[i] = run;
        run += c;
        count[i] = 0;  // cleared for the next rebuild, so the counts are zero whenever a rebuild starts
    }
    if (t == 0) start[ncell] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(int ncell, const int *__restrict__ count,
                                                              int *__restrict__ tileSum) {
    __shared__ int waveTot[SCAN_THREADS / WAVE];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int sum = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) sum += base + k < ncell ? count[base + k] : 0;
    int total = 0;
    block_exclusive_scan(sum, waveTot, &total);
    if (threadIdx.x == 0) tileSum[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tile_sums(int tiles, int *__restrict__ tileSum) {
    __shared__ int waveTot[SCAN_THREADS / WAVE];
    const int t = threadIdx.x;
    const int v = t < tiles ? tileSum[t] : 0;
    int total = 0;
    const int excl = block_exclusive_scan(v, waveTot, &total);
    if (t < tiles) tileSum[t] = excl;
    if (t == 0) tileSum[tiles] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(int ncell, int tiles, int *__restrict__ count,
                                                              const int *__restrict__ tileSum,
                                                              int *__restrict__ start) {
    __shared__ int waveTot[SCAN_THREADS / WAVE];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int c[SCAN_ITEMS];
    int sum = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        c[k] = base + k < ncell ? count[base + k] : 0;
        sum += c[k];
    }
    int total = 0;
    int run = block_exclusive_scan(sum, waveTot, &total) + tileSum[blockIdx.x];
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < ncell) {
            start[base + k] = run;
            count[base + k] = 0;
        }
        run += c[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) start[ncell] = tileSum[tiles];
}

__global__ __launch_bounds__(256) void k_scatter(int n, const int *__restrict__ cellStart,
                                                 const int *__restrict__ tmpCell, const int *__restrict__ tmpSlot,
                                                 int *__restrict__ tmpOrder) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tmpOrder[cellStart[tmpCell[i]] + tmpSlot[i]] = i;
}

// One thread per source record: its rank among its cell's members by source index (a stable order), then its record
// into the destination arrays at the cell's start plus that rank.  A grid of one cell holds every record in source
// order.
__global__ __launch_bounds__(256) void k_gather(Grid g, int n, Particles src, Particles dst,
                                                const double *__restrict__ tmpPos, const uint8_t *__restrict__ tmpW,
                                                const int *__restrict__ tmpOrder, const int *__restrict__ tmpCell) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int cell = tmpCell[i];
    const int start = g.cellStart[cell];
    const int m = g.cellStart[cell + 1] - start;
    int rank = 0;
    if (m == n) {
        rank = i;
    } else {
        for (int q = 0; q < m; q++) rank += tmpOrder[start + q] < i ? 1 : 0;
    }
    const int d = start + rank;
    dst.x[d] = tmpPos[i];
    dst.y[d] = tmpPos[n + i];
    dst.z[d] = tmpPos[2 * n + i];
    for (int c = 0; c < 9; c++) dst.o[c][d] = src.o[c][i];
    dst.tag[d] = src.tag[i];
    dst.wbits[d] = tmpW[i];
    dst.cell[d] = cell;
}

size_t rebuild_scan_ints(int ncell) {
    return ncell > SCAN_ONE_BLOCK ? (size_t)((ncell + SCAN_TILE - 1) / SCAN_TILE + 1) : 0;
}

hipError_t launch_rebuild(hipStream_t s, const Grid &g, int n, Particles src, Particles dst, int *tmpSlot,
                          int *tmpOrder, double *tmpPos, int rescale, const double *oldInv, Timed t, int *scanTmp) {
    // tmpSlot holds 2n ints (slot, cell); tmpOrder holds n ints followed by n bytes of wrap bits; scanTmp holds
    // rebuild_scan_ints(ncell) ints.  The cell counts are zero here: cleared when allocated and by every scan.
    int *tmpCell = tmpSlot + n;
    uint8_t *tmpW = reinterpret_cast<uint8_t *>(tmpOrder + n);
    M3 inv{};
    if (rescale)
        for (int i = 0; i < 9; i++) inv.m[i] = oldInv[i];
    unsigned blocks = (unsigned)((n + 255) / 256);
    hipExtLaunchKernelGGL(k_bin, dim3(blocks), dim3(256), 0, s, t.start, nullptr, 0, g, n, src, rescale, inv, tmpPos,
                          tmpCell, tmpSlot, tmpW);
    if (g.ncell <= SCAN_ONE_BLOCK) {
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(SCAN_THREADS), 0, s, g.ncell, g.cellCount, g.cellStart);
    } else {
        const int tiles = (g.ncell + SCAN_TILE - 1) / SCAN_TILE;
        if (tiles > SCAN_MAX_TILES || !scanTmp) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(SCAN_THREADS), 0, s, g.ncell, g.cellCount, scanTmp);
        hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(SCAN_THREADS), 0, s, tiles, scanTmp);
        hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(SCAN_THREADS), 0, s, g.ncell, tiles, g.cellCount, scanTmp,
                           g.cellStart);
    }
    hipLaunchKernelGGL(k_scatter, dim3(blocks), dim3(256), 0, s, n, g.cellStart, tmpCell, tmpSlot, tmpOrder);
    hipExtLaunchKernelGGL(k_gather, dim3(blocks), dim3(256), 0, s, nullptr, t.stop, 0, g, n, src, dst, tmpPos, tmpW,
                          tmpOrder, tmpCell);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------------------
// Shared pieces of the stencil kernels

// Stencil view of one block: entries of up to 27 stencil cells (own cell first) either staged in LDS or read from
// the cell-sorted arrays in HBM when they do not fit the LDS area.
struct Stencil {
    const int *start;   // prefix offsets of the stencil cells (LDS)
    const int *base;    // first record index of each stencil cell (LDS)
    const int *ecode;   // packed image shift of each stencil cell (LDS)
    int ncells;
    double *x, *y, *z, *ax, *ay, *az;
    int *code;
};

__device__ __forceinline__ int stencil_cell(const int *start, int ncells, int j) {
    int lo = 0, hi = ncells;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (start[mid] <= j) lo = mid; else hi = mid;
    }
    return lo;
}

// Per-axis image code e + w_j + 1 (two bits per axis) of a record with wrap bits w in a stencil cell with packed
// image shift ec (e + 1 per axis).
__device__ __forceinline__ int image_code(int ec, int w) {
    return (((ec & 3) - 1 + (w & 1)) + 1) | ((((ec >> 2) & 3) - 1 + ((w >> 1) & 1) + 1) << 2) |
           ((((ec >> 4) & 3) - 1 + ((w >> 2) & 1) + 1) << 4);
}

// Translation box * (code - 1 - w) from the LDS copy of the reference's 27-entry translation table.
__device__ __forceinline__ V3 image_translation(const double *sT, int code, int wx, int wy, int wz) {
    int kx = (code & 3) - 1 - wx, ky = ((code >> 2) & 3) - 1 - wy, kz = ((code >> 4) & 3) - 1 - wz;
    int t = 3 * ((kx + 1) * 9 + (ky + 1) * 3 + (kz + 1));
    return v3(sT[t], sT[t + 1], sT[t + 2]);
}

template <int KIND, bool STAGED>
__device__ __forceinline__ void load_candidate(const Particles &p, const Stencil &S, int j, V3 &p2, V3 &a2,
                                               int &code) {
    if (STAGED) {
        p2 = v3(S.x[j], S.y[j], S.z[j]);
        a2 = KIND == 1 ? v3(S.ax[j], S.ay[j], S.az[j]) : v3(0, 0, 0);
        code = S.code[j];
    } else {
        int q = stencil_cell(S.start, S.ncells, j);
        int gi = S.base[q] + (j - S.start[q]);
        p2 = load3(p.x, p.y, p.z, gi);
        a2 = KIND == 1 ? v3(p.o[2][gi], p.o[5][gi], p.o[8][gi]) : v3(0, 0, 0);
        code = image_code(S.ecode[q], p.wbits[gi]);
    }
}

constexpr int MAX_STENCIL = 125;  // (2 * reach + 1)^3 cells for reach <= 2
constexpr int MAX_HALF = 63;      // own cell + the positive half of the other 124

// Number of stencil cells: the full (2k+1)^3 block, or the own cell plus the lexicographically positive half.
__device__ __forceinline__ int stencil_cells(int reach, bool half) {
    const int side = 2 * reach + 1, all = side * side * side;
    return half ? (all - 1) / 2 + 1 : all;
}

// Fills one slot of the stencil cell table of cell (cx, cy, cz): slot 0 is the own cell, then either all other
// offsets in [-k, k]^3 (sweep) or only the positive ones, (o0, o1, o2) > 0 lexicographically with o0 the most
// significant -- for k = 1 the reference's 9*n0 + 3*n1 + n2 > 13 (NeighbourGrid.cpp:130-131).
// Reach-2 stencil cells (index (dx + 2) * 25 + (dy + 2) * 5 + dz + 2) nearest first: the own cell, the 26 cells one
// step away (Chebyshev distance), then the 98 two steps away, each shell in natural order.  Read-only.
__constant__ unsigned char kShellOrder[125] = {
    62, 31, 32, 33, 36, 37, 38, 41, 42, 43, 56, 57, 58, 61, 63, 66, 67, 68, 81, 82, 83, 86, 87, 88, 91,
    92, 93, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22,
    23, 24, 25, 26, 27, 28, 29, 30, 34, 35, 39, 40, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53, 54, 55, 59,
    60, 64, 65, 69, 70, 71, 72, 73, 74, 75, 76, 77, 78, 79, 80, 84, 85, 89, 90, 94, 95, 96, 97, 98, 99,
    100, 101, 102, 103, 104, 105, 106, 107, 108, 109, 110, 111, 112, 113, 114, 115, 116, 117, 118, 119, 120, 121, 122, 123, 124
};

__device__ __forceinline__ void stencil_table(const Grid &g, int cx, int cy, int cz, int slot, bool half,
                                              int *s_start, int *s_base, int *s_ecode, bool byShell = false) {
    const int k = g.reach, side = 2 * k + 1, centre = (side * side * side - 1) / 2;
    int t;
    if (slot == 0) t = centre;
    else if (half) t = centre + slot;
    else if (byShell && k == 2) t = kShellOrder[slot];
    else t = slot <= centre ? slot - 1 : slot;
    int ex, ey, ez;
    int nx = wrap_axis(cx, t / (side * side) - k, g.n[0], ex);
    int ny = wrap_axis(cy, (t / side) % side - k, g.n[1], ey);
    int nz = wrap_axis(cz, t % side - k, g.n[2], ez);
    int nc = cell_index(g, nx, ny, nz);
    s_base[slot] = g.cellStart[nc];
    s_start[slot + 1] = g.cellStart[nc + 1] - g.cellStart[nc];
    s_ecode[slot] = (ex + 1) | ((ey + 1) << 2) | ((ez + 1) << 4);
}

// Builds the table of `ncells` stencil cells and its prefix sums (s_start[ncells] = total entries).
// byShell (full stencils of reach 2 only) orders the cells by distance from the own cell, so that a search that
// stops at the first overlap meets the nearest candidates first; the set of cells is the same.
__device__ __forceinline__ void build_stencil(const Grid &g, int cx, int cy, int cz, int ncells, bool half,
                                              int *s_start, int *s_base, int *s_ecode, bool byShell = false) {
    for (int slot = threadIdx.x; slot < ncells; slot += blockDim.x)
        stencil_table(g, cx, cy, cz, slot, half, s_start, s_base, s_ecode, byShell);
    __syncthreads();
    if (threadIdx.x == 0) {
        s_start[0] = 0;
        for (int q = 1; q <= ncells; q++) s_start[q] += s_start[q - 1];
    }
    __syncthreads();
}

template <int KIND>
__device__ __forceinline__ void stage_stencil(const Particles &p, const Stencil &S, int total, int tid, int nt) {
    for (int j = tid; j < total; j += nt) {
        int q = stencil_cell(S.start, S.ncells, j);
        int gi = S.base[q] + (j - S.start[q]);
        S.x[j] = p.x[gi];
        S.y[j] = p.y[gi];
        S.z[j] = p.z[gi];
        if (KIND == 1) {
            S.ax[j] = p.o[2][gi];
            S.ay[j] = p.o[5][gi];
            S.az[j] = p.o[8][gi];
        }
        if (KIND == 2) S.ax[j] = (double)p.tag[gi];  // bead item index: particle * nc + bead
        S.code[j] = image_code(S.ecode[q], p.wbits[gi]);
    }
}

__device__ __forceinline__ void carve_stencil(Stencil &S, double *smem, int cap, int kind, double *&rest) {
    S.x = smem;
    S.y = S.x + cap;
    S.z = S.y + cap;
    S.ax = S.z + cap;
    S.ay = S.ax + cap;
    S.az = S.ay + cap;
    double *after = kind == 0 ? S.ax : S.az + cap;  // kind 2 keeps the bead item index in ax
    S.code = reinterpret_cast<int *>(after);
    rest = after + (cap + 1) / 2;
}

size_t stencil_lds_bytes(int kind, int cap) {
    return (size_t)(kind == 0 ? 3 : 6) * sizeof(double) * cap + sizeof(double) * ((cap + 1) / 2);
}

// ---------------------------------------------------------------------------------------------------------------
// Half-shell overlap count (Packing::countTotalOverlapsNGCellHelper, Packing.cpp:431-517): one workgroup per cell
// stages the cell and its 13 positive neighbours in LDS.  Within a cell the particle with the higher caller index
// is the first operand, as in the reference's LIFO cell lists; across cells the first operand lies in the base
// cell.

constexpr int COUNT_THREADS = 256;

// The early-exit flag of an overlap count (counters[CNT_OVERLAPS] != 0) read at device scope: each XCD has its own
// L2, and a plain load can keep returning a line cached there before another workgroup counted an overlap (the
// count itself is a device-scope atomic), so workgroups elsewhere would not stop.
__device__ __forceinline__ bool overlap_found(unsigned long long *counters) {
    return __hip_atomic_load(counters + CNT_OVERLAPS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

template <int KIND, bool STAGED>
__device__ void count_cell(const Grid &g, const Particles &p, const PairConsts &k, const Stencil &S, const double *sT,
                           const int *sTag, int nOwn, int total, int earlyExit, unsigned long long *counters,
                           int *sFound) {
    const int tid = threadIdx.x;
    unsigned long long found = 0;
    for (int a = 0; a < nOwn; a++) {
        V3 pos1, ax1;
        int c1;
        load_candidate<KIND, STAGED>(p, S, a, pos1, ax1, c1);
        const int wx = (c1 & 3) - 1, wy = ((c1 >> 2) & 3) - 1, wz = ((c1 >> 4) & 3) - 1;
        const int tagA = sTag[a];
        SCFirst f = make_first<KIND>(k, pos1, ax1);
        for (int b = tid; b < total; b += COUNT_THREADS) {
            if (b < nOwn && sTag[b] >= tagA) continue;
            V3 p2, a2;
            int code;
            load_candidate<KIND, STAGED>(p, S, b, p2, a2, code);
            V3 T = b < nOwn && code == c1 ? v3(0, 0, 0) : image_translation(sT, code, wx, wy, wz);
            if (pair_overlap<KIND>(k, f, p2, a2, T)) found++;
        }
        if (earlyExit) {
            int any = __syncthreads_or(found != 0 || overlap_found(counters));
            if (any) break;
        }
    }
    if (found) atomicAdd(sFound, (int)found);
    __syncthreads();
    if (tid == 0 && *sFound) atomicAdd(&counters[CNT_OVERLAPS], (unsigned long long)*sFound);
}

// Flat form of count_cell for a staged stencil: the (own particle, candidate) pairs of the cell are processed in
// chunks of COUNT_CHUNK by all threads at once -- pass A classifies each pair by its centre distance (definite
// overlap, ambiguous, far), pass B runs the segment test on the chunk's ambiguous pairs, compacted block-wide -- so
// no thread walks the own particles one by one.  Same arithmetic as pair_overlap, split in two.
constexpr int COUNT_OWN_CAP = 128;             // own particles whose first-operand data fits the table
constexpr int COUNT_CHUNK = COUNT_THREADS * 8;  // pairs per chunk (and queue entries)

// KIND 2 (bead items of polyspheres): pairs of beads of one particle are skipped (Packing.cpp:482-483, 500-501)
// and the contact distance is the sum of the two bead radii (sBeadR, indexed by bead).
template <int KIND>
__device__ void count_cell_flat(const PairConsts &k, const Stencil &S, const double *sT, const int *sTag, int nOwn,
                                int total, int earlyExit, unsigned long long *counters, int *sFound, double *fTab,
                                int *queue, int *sQueued, const double *sBeadR) {
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    // first operands: position, cap A, u, a (10 doubles) per own particle
    for (int a = tid; a < nOwn; a += COUNT_THREADS) {
        const V3 pos1 = v3(S.x[a], S.y[a], S.z[a]);
        SCFirst f = make_first<KIND>(k, pos1, KIND == 1 ? v3(S.ax[a], S.ay[a], S.az[a]) : v3(0, 0, 0));
        double *q = fTab + 10 * a;
        q[0] = f.pos.x; q[1] = f.pos.y; q[2] = f.pos.z;
        if (KIND == 1) {
            q[3] = f.capA.x; q[4] = f.capA.y; q[5] = f.capA.z;
            q[6] = f.u.x; q[7] = f.u.y; q[8] = f.u.z;
            q[9] = f.a;
        }
    }
    if (tid == 0) *sQueued = 0;
    __syncthreads();
    unsigned long long found = 0;
    const int pairs = nOwn * total;  // < COUNT_OWN_CAP * 65536
    // pair index q = a * total + b, advanced by COUNT_THREADS per step without a division
    int qa = tid / total, qb = tid - (tid / total) * total;
    for (int base = 0; base < pairs; base += COUNT_CHUNK) {
        // pass A
        for (int i = 0; i < COUNT_CHUNK / COUNT_THREADS; i++) {
            const int qi = base + i * COUNT_THREADS + tid;
            bool amb = false;
            const int a = qa, b = qb;
            qb += COUNT_THREADS;
            while (qb >= total) {
                qb -= total;
                qa++;
            }
            if (qi < pairs) {
                if (!(b < nOwn && sTag[b] >= sTag[a])) {
                    const int c1 = S.code[a], code = S.code[b];
                    const int wx = (c1 & 3) - 1, wy = ((c1 >> 2) & 3) - 1, wz = ((c1 >> 4) & 3) - 1;
                    const V3 T = b < nOwn && code == c1 ? v3(0, 0, 0) : image_translation(sT, code, wx, wy, wz);
                    const double *q = fTab + 10 * a;
                    if (KIND == 2) {
                        const int ta = (int)S.ax[a], tb = (int)S.ax[b];
                        const int pa = ta / k.nc, pb = tb / k.nc;
                        if (pa != pb &&
                            bead_overlap(sBeadR[ta - pa * k.nc], sBeadR[tb - pb * k.nc], v3(q[0], q[1], q[2]),
                                         v3(S.x[b], S.y[b], S.z[b]), T))
                            found++;
                    } else {
                        const V3 d = sub(add(v3(S.x[b], S.y[b], S.z[b]), T), v3(q[0], q[1], q[2]));
                        const double d2 = dot(d, d);
                        if (d2 < k.contact2) found++;
                        else if (KIND == 1 && d2 < k.range2) amb = true;
                    }
                }
            }
            if (KIND == 1) {
                const unsigned long long m = __ballot(amb);
                int slot = 0;
                if (lane == 0 && m) slot = atomicAdd(sQueued, __popcll(m));
                slot = __shfl(slot, 0);
                if (amb) queue[slot + __popcll(m & ((1ull << lane) - 1))] = (a << 16) | b;
            }
        }
        __syncthreads();
        // pass B
        if (KIND == 1) {
            const int queued = *sQueued;
            for (int e = tid; e < queued; e += COUNT_THREADS) {
                const int a = queue[e] >> 16, b = queue[e] & 0xffff;
                const int c1 = S.code[a], code = S.code[b];
                const int wx = (c1 & 3) - 1, wy = ((c1 >> 2) & 3) - 1, wz = ((c1 >> 4) & 3) - 1;
                const V3 T = b < nOwn && code == c1 ? v3(0, 0, 0) : image_translation(sT, code, wx, wy, wz);
                const double *q = fTab + 10 * a;
                const V3 p2 = add(v3(S.x[b], S.y[b], S.z[b]), T);
                const V3 a2 = v3(S.ax[b], S.ay[b], S.az[b]);
                const V3 c2a = add(p2, scal(a2, -k.halfLen));
                const V3 c2b = add(p2, scal(a2, k.halfLen));
                if (segment_distance2(v3(q[3], q[4], q[5]), v3(q[6], q[7], q[8]), q[9], c2a, c2b) < k.contact2)
                    found++;
            }
        }
        const int stop = earlyExit ? __syncthreads_or(found != 0 || overlap_found(counters)) : (__syncthreads(), 0);
        if (tid == 0) *sQueued = 0;
        __syncthreads();
        if (stop) break;
    }
    if (found) atomicAdd(sFound, (int)found);
    __syncthreads();
    if (tid == 0 && *sFound) atomicAdd(&counters[CNT_OVERLAPS], (unsigned long long)*sFound);
}

// rot: workgroup b counts cell (b + rot) mod ncell, so that an early-exit count starts where the previous one found
// an overlap (counters[CNT_SCAN_HINT], written by a workgroup that finds one); the count does not depend on it.
template <int KIND>
__global__ __launch_bounds__(COUNT_THREADS) void k_count_overlaps(Grid g, Particles p, PairConsts k, int earlyExit,
                                                                  unsigned long long *counters, int cap, int rot) {
    extern __shared__ double smem[];
    __shared__ int s_start[MAX_HALF + 1], s_base[MAX_HALF], s_ecode[MAX_HALF];
    __shared__ int sFound;
    __shared__ double sT[81];
    const int tid = threadIdx.x;
    const int cell = (int)((blockIdx.x + (unsigned)rot) % (unsigned)g.ncell);
    const int own0 = g.cellStart[cell];
    const int nOwn = g.cellStart[cell + 1] - own0;
    if (nOwn == 0) return;
    if (earlyExit && overlap_found(counters)) return;
    const int cx = cell % g.n[0], cy = (cell / g.n[0]) % g.n[1], cz = cell / (g.n[0] * g.n[1]);
    const int ncells = stencil_cells(g.reach, true);
    for (int i = tid; i < 81; i += COUNT_THREADS) sT[i] = g.T[i];
    if (tid == 0) sFound = 0;
    build_stencil(g, cx, cy, cz, ncells, true, s_start, s_base, s_ecode);
    const int total = s_start[ncells];
    Stencil S;
    S.start = s_start;
    S.base = s_base;
    S.ecode = s_ecode;
    S.ncells = ncells;
    double *rest;
    carve_stencil(S, smem, cap, KIND, rest);
    int *sTag = reinterpret_cast<int *>(rest);  // own-cell tags, up to cap entries
    for (int a = tid; a < nOwn && a < cap; a += COUNT_THREADS) sTag[a] = p.tag[own0 + a];
    __shared__ double sBeadR[RP_MAX_CENTRES];
    if (KIND == 2)
        for (int c = tid; c < k.nc; c += COUNT_THREADS) sBeadR[c] = k.beads[4 * c + 3];
    if (total <= cap && nOwn <= COUNT_OWN_CAP && total <= 0xffff) {
        __shared__ int sQueued;
        stage_stencil<KIND>(p, S, total, tid, COUNT_THREADS);
        double *fTab = reinterpret_cast<double *>(sTag + ((cap + 1) & ~1));  // 10 doubles per own particle
        int *queue = reinterpret_cast<int *>(fTab + 10 * COUNT_OWN_CAP);        // COUNT_CHUNK entries
        __syncthreads();
        count_cell_flat<KIND>(k, S, sT, sTag, nOwn, total, earlyExit, counters, &sFound, fTab, queue, &sQueued,
                              sBeadR);
    } else if (KIND == 2) {
        // bead grids hold a few beads per cell; a cell beyond the flat path's limits is reported as an error
        if (tid == 0) atomicAdd(&counters[CNT_STAGE_OVERFLOW], 1ull);
    } else if (total <= cap) {
        stage_stencil<KIND>(p, S, total, tid, COUNT_THREADS);
        __syncthreads();
        count_cell<KIND, true>(g, p, k, S, sT, sTag, nOwn, total, earlyExit, counters, &sFound);
    } else if (nOwn <= cap) {
        __syncthreads();
        count_cell<KIND, false>(g, p, k, S, sT, sTag, nOwn, total, earlyExit, counters, &sFound);
    } else if (tid == 0) {
        // a cell holding more than `cap` particles: the host sizes cap from the largest cell, so this marks an error
        atomicAdd(&counters[CNT_STAGE_OVERFLOW], 1ull);
    }
    if (tid == 0 && sFound) counters[CNT_SCAN_HINT] = (unsigned long long)cell;
}

size_t count_lds_bytes(int kind, int cap) {
    return stencil_lds_bytes(kind, cap) + sizeof(int) * (((cap + 1) & ~1) + COUNT_CHUNK) +
           sizeof(double) * 10 * COUNT_OWN_CAP;
}

hipError_t launch_count_overlaps(hipStream_t s, const Grid &g, int n, Particles p, PairConsts k, int earlyExit,
                                 unsigned long long *counters, int cap, Timed t, int startCell) {
    if (n == 0) return hipSuccess;
    size_t lds = count_lds_bytes(k.kind, cap);
    const int rot = g.ncell > 0 ? ((startCell % g.ncell) + g.ncell) % g.ncell : 0;
    if (k.kind == 0)
        hipExtLaunchKernelGGL(k_count_overlaps<0>, dim3((unsigned)g.ncell), dim3(COUNT_THREADS), (uint32_t)lds, s,
                              t.start, t.stop, 0, g, p, k, earlyExit, counters, cap, rot);
    else if (k.kind == 1)
        hipExtLaunchKernelGGL(k_count_overlaps<1>, dim3((unsigned)g.ncell), dim3(COUNT_THREADS), (uint32_t)lds, s,
                              t.start, t.stop, 0, g, p, k, earlyExit, counters, cap, rot);
    else
        hipExtLaunchKernelGGL(k_count_overlaps<2>, dim3((unsigned)g.ncell), dim3(COUNT_THREADS), (uint32_t)lds, s,
                              t.start, t.stop, 0, g, p, k, earlyExit, counters, cap, rot);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------------------
// Total soft energy (Packing::getTotalEnergy / getTotalEnergyNGCellHelper, Packing.cpp:690-709, 766-839): the same
// half-shell traversal as the overlap count, every pair once; one partial per cell, then a fixed-order sum.

constexpr int SUM_THREADS = 256;
constexpr int SUM_BLOCKS = 256;

// Sums SUM_THREADS values held one per thread in a fixed tree; the result is in lds[0].
__device__ __forceinline__ double block_sum(double v, double *lds) {
    const int tid = threadIdx.x;
    lds[tid] = v;
    __syncthreads();
    for (int w = SUM_THREADS / 2; w > 0; w >>= 1) {
        if (tid < w) lds[tid] += lds[tid + w];
        __syncthreads();
    }
    return lds[0];
}

template <int KIND, bool STAGED>
__device__ double energy_cell(const Particles &p, const PairConsts &k, const Stencil &S, const double *sT,
                              const int *sTag, int nOwn, int total) {
    const int tid = threadIdx.x;
    double e = 0;
    for (int a = 0; a < nOwn; a++) {
        V3 pos1, ax1;
        int c1;
        load_candidate<KIND, STAGED>(p, S, a, pos1, ax1, c1);
        const int wx = (c1 & 3) - 1, wy = ((c1 >> 2) & 3) - 1, wz = ((c1 >> 4) & 3) - 1;
        const int tagA = sTag[a];
        SCFirst f = make_first<KIND>(k, pos1, ax1);
        for (int b = tid; b < total; b += COUNT_THREADS) {
            if (b < nOwn && sTag[b] >= tagA) continue;
            V3 p2, a2;
            int code;
            load_candidate<KIND, STAGED>(p, S, b, p2, a2, code);
            V3 T = b < nOwn && code == c1 ? v3(0, 0, 0) : image_translation(sT, code, wx, wy, wz);
            e += soft_energy(k, f, p2, a2, T);
        }
    }
    return e;
}

template <int KIND>
__global__ __launch_bounds__(COUNT_THREADS) void k_total_energy(Grid g, Particles p, PairConsts k, double *cellEnergy,
                                                                unsigned long long *counters, int cap) {
    extern __shared__ double smem[];
    __shared__ int s_start[MAX_HALF + 1], s_base[MAX_HALF], s_ecode[MAX_HALF];
    __shared__ double sT[81];
    __shared__ double sSum[COUNT_THREADS];
    const int tid = threadIdx.x;
    const int cell = blockIdx.x;
    const int own0 = g.cellStart[cell];
    const int nOwn = g.cellStart[cell + 1] - own0;
    if (nOwn == 0) {
        if (tid == 0) cellEnergy[cell] = 0;
        return;
    }
    const int cx = cell % g.n[0], cy = (cell / g.n[0]) % g.n[1], cz = cell / (g.n[0] * g.n[1]);
    const int ncells = stencil_cells(g.reach, true);
    for (int i = tid; i < 81; i += COUNT_THREADS) sT[i] = g.T[i];
    build_stencil(g, cx, cy, cz, ncells, true, s_start, s_base, s_ecode);
    const int total = s_start[ncells];
    Stencil S;
    S.start = s_start;
    S.base = s_base;
    S.ecode = s_ecode;
    S.ncells = ncells;
    double *rest;
    carve_stencil(S, smem, cap, KIND, rest);
    int *sTag = reinterpret_cast<int *>(rest);
    for (int a = tid; a < nOwn && a < cap; a += COUNT_THREADS) sTag[a] = p.tag[own0 + a];
    double e = 0;
    if (total <= cap) {
        stage_stencil<KIND>(p, S, total, tid, COUNT_THREADS);
        __syncthreads();
        e = energy_cell<KIND, true>(p, k, S, sT, sTag, nOwn, total);
    } else if (nOwn <= cap) {
        __syncthreads();
        e = energy_cell<KIND, false>(p, k, S, sT, sTag, nOwn, total);
    } else if (tid == 0) {
        atomicAdd(&counters[CNT_STAGE_OVERFLOW], 1ull);
    }
    static_assert(COUNT_THREADS == SUM_THREADS, "one partial per thread");
    double sum = block_sum(e, sSum);
    if (tid == 0) cellEnergy[cell] = sum;
}

// Fixed-order sum of n doubles: SUM_BLOCKS strided partials, then one workgroup; out = scratch[SUM_BLOCKS].
__global__ __launch_bounds__(SUM_THREADS) void k_sum_partial(const double *in, int n, double *scratch) {
    __shared__ double lds[SUM_THREADS];
    double v = 0;
    for (int i = blockIdx.x * SUM_THREADS + threadIdx.x; i < n; i += SUM_BLOCKS * SUM_THREADS) v += in[i];
    double s = block_sum(v, lds);
    if (threadIdx.x == 0) scratch[blockIdx.x] = s;
}

__global__ __launch_bounds__(SUM_THREADS) void k_sum_final(double *scratch) {
    __shared__ double lds[SUM_THREADS];
    static_assert(SUM_BLOCKS == SUM_THREADS, "one partial per thread");
    double s = block_sum(scratch[threadIdx.x], lds);
    if (threadIdx.x == 0) scratch[SUM_BLOCKS] = s;
}

size_t fixed_sum_scratch_doubles() { return SUM_BLOCKS + 1; }

hipError_t launch_total_energy(hipStream_t s, const Grid &g, int n, Particles p, PairConsts k, double *cellEnergy,
                               double *scratch, unsigned long long *counters, int cap) {
    size_t lds = count_lds_bytes(k.kind, cap);
    if (n > 0 && g.ncell > 0) {
        if (k.kind == 0)
            hipLaunchKernelGGL(k_total_energy<0>, dim3((unsigned)g.ncell), dim3(COUNT_THREADS), lds, s, g, p, k,
                               cellEnergy, counters, cap);
        else
            hipLaunchKernelGGL(k_total_energy<1>, dim3((unsigned)g.ncell), dim3(COUNT_THREADS), lds, s, g, p, k,
                               cellEnergy, counters, cap);
    }
    // with no particles the partials are sums of nothing (zeros)
    hipLaunchKernelGGL(k_sum_partial, dim3(SUM_BLOCKS), dim3(SUM_THREADS), 0, s, cellEnergy, n > 0 ? g.ncell : 0,
                       scratch);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(SUM_THREADS), 0, s, scratch);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------------------
// Checkerboard sweep.  Workgroup = one active cell of colour `colour` (cells with coordinates of the colour's
// parity on every axis; same-colour cells are separated by a full inactive cell, wider than the interaction
// range).  The cell's particles are moved n_cell times: uniform random pick + rototranslation
// (RototranslationSampler.cpp:22-49).  Proposals are drawn in parallel (one lane per move, its own Philox
// stream), then applied in order; each trial pose is tested against the 27-cell stencil 64*WAVES candidates per
// step with a workgroup vote after every step, and the first step with an overlap rejects the move.

constexpr int OWN_CAP = 64;   // own-cell particles kept in LDS
constexpr int SWEEP_AUX = 288;   // doubles of the planner / soft scratch area (grouped: 3 buffers x 4 plans x 24)
constexpr int SWEEP_AUX_I = 32;  // ints of planner output (grouped: 3 buffers x 8, plus 3 hit words)
constexpr int GROUPED_QUEUE = 512;  // queued (candidate, move) pairs per tester wavefront in the grouped sweep
constexpr int MOVE_CHUNK = 64; // proposals staged per batch
constexpr int SLOT_GROUP = 4;  // candidate slots loaded per lane before any is tested

// Proposals of one cycle for every (cell, move): drawn in one fully parallel launch before the colour passes, one
// Philox stream per (cycle, cell, move) -- RototranslationSampler::sampleMove (RototranslationSampler.cpp:22-49):
// particle pick among the cell's particles, translation in [-t, t)^3, axis by rejection in the unit ball, angle in
// [-r, r), rotation by Rodrigues' formula (Matrix.tpp:438-450).
__device__ __forceinline__ void rodrigues(V3 axis, double angle, double *R) {
    double nrm = sqrt(dot(axis, axis));
    V3 ua = v3(axis.x / nrm, axis.y / nrm, axis.z / nrm);
    double sa = sin(angle), ca = cos(angle);
    double K[9] = {0, -ua.z, ua.y, ua.z, 0, -ua.x, -ua.y, ua.x, 0};
    double f1 = 1 - ca;
    double fK[9];
    for (int c = 0; c < 9; c++) fK[c] = K[c] * f1;
    double fKK[9];
    mulm(fK, K, fKK);
    for (int c = 0; c < 9; c++) R[c] = ((c % 4 == 0 ? 1.0 : 0.0) + K[c] * sa) + fKK[c];
}

// The move arithmetic of the sweep on explicit inputs, one move per thread: the trial position (translate_wrapped), the
// rotation matrix of (axis, angle) (rodrigues) and the trial orientation R O (mulm), the functions k_sweep and
// k_proposals use.  Arrays of n records: pos 3, rot 9, t 3, axis 3, angle 1; outputs pos 3, R 9, rot 9, sin/cos 2.
__global__ __launch_bounds__(256) void k_apply_moves(M3 dims, M3 inv, size_t n, const double *__restrict__ pos,
                                                     const double *__restrict__ rot, const double *__restrict__ t,
                                                     const double *__restrict__ axis, const double *__restrict__ angle,
                                                     double *outPos, double *outR, double *outRot, double *outSc) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V3 p = translate_wrapped(dims.m, inv.m, v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
                                   v3(t[3 * i], t[3 * i + 1], t[3 * i + 2]));
    outPos[3 * i] = p.x;
    outPos[3 * i + 1] = p.y;
    outPos[3 * i + 2] = p.z;
    double R[9], O[9], NO[9];
    rodrigues(v3(axis[3 * i], axis[3 * i + 1], axis[3 * i + 2]), angle[i], R);
    for (int c = 0; c < 9; c++) O[c] = rot[9 * i + c];
    mulm(R, O, NO);
    for (int c = 0; c < 9; c++) {
        outR[9 * i + c] = R[c];
        outRot[9 * i + c] = NO[c];
    }
    outSc[2 * i] = sin(angle[i]);
    outSc[2 * i + 1] = cos(angle[i]);
}

// Folds the move-count slot of cell i (one lane per cell, every lane of the wavefront calling) into the counters
// and clears it: wavefront sums, one atomic per word and wavefront.
__device__ __forceinline__ void fold_slots_wave(const FoldSlots &f, int i) {
    unsigned long long v[SLOT_WORDS];
    for (int w = 0; w < SLOT_WORDS; w++) v[w] = 0;
    if (i < f.cells) {
        unsigned long long *slot = f.slots + (size_t)i * SLOT_WORDS;
        for (int w = 0; w < SLOT_WORDS; w++) {
            v[w] = slot[w];
            slot[w] = 0;
        }
    }
    for (int w = 0; w < SLOT_WORDS; w++)
        for (int off = WAVE / 2; off > 0; off >>= 1) v[w] += __shfl_xor(v[w], off);
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        const int target[SLOT_WORDS] = {CNT_ACCEPTED, CNT_ATTEMPTED, CNT_LEFT_CELL, CNT_TESTS, CNT_G32_PAIRS};
        for (int w = 0; w < SLOT_WORDS; w++)
            if (v[w]) atomicAdd(&f.counters[target[w]], v[w]);
    }
}

__global__ __launch_bounds__(256) void k_fold_slots(FoldSlots f) {
    fold_slots_wave(f, blockIdx.x * blockDim.x + threadIdx.x);
}

hipError_t launch_fold_slots(hipStream_t s, unsigned long long *slots, int ncell, unsigned long long *counters) {
    if (ncell <= 0) return hipSuccess;
    FoldSlots f{slots, ncell, counters};
    hipLaunchKernelGGL(k_fold_slots, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s, f);
    return hipGetLastError();
}

// moves > 0: `moves` proposals per cell at [cell * moves, ...); moves == 0: ms.n per particle, cell c's at
// [ms.n cellStart[c], ms.n cellStart[c + 1]) (the caller passes the cell of every sorted record in `cellOf`).
// With several samplers the first draw picks the sampler (Simulation.cpp:456-465); a translation sampler's move
// has the identity rotation and a rotation sampler's a zero translation, which leave the pose bit-identical.
// fold.slots: the previous cycle's move-count slots are folded into the counters first (fold_slots_wave; the grid
// covers fold.cells threads as well).
__global__ __launch_bounds__(256) void k_proposals(Grid g, Proposals pr, int moves, const int *__restrict__ cellOf,
                                                   int n, MoveSet ms, uint64_t seed, uint32_t cycle, FoldSlots fold) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (fold.slots) fold_slots_wave(fold, i);
    int cell, m;
    if (moves > 0) {
        if (i >= g.ncell * moves) return;
        cell = i / moves;
        m = i - cell * moves;
    } else {
        if (i >= n * ms.n) return;
        cell = cellOf[i / ms.n];
        m = i - g.cellStart[cell] * ms.n;
    }
    const int nOwn = g.cellStart[cell + 1] - g.cellStart[cell];
    if (nOwn == 0) {
        if (pr.type) pr.type[i] = 255;
        if (pr.acc) pr.acc[i] = 0;
        return;
    }
    Philox rng = philox_init(seed, cycle, (uint32_t)cell, (uint32_t)m);
    const int type = ms.n > 1 ? (int)philox_below(rng, (uint32_t)ms.n) : 0;
    const int kind = ms.kind[type];
    const double ts = ms.ts[type], rs = ms.rs[type];
    pr.k[i] = (int)philox_below(rng, (uint32_t)nOwn);
    for (int d = 0; d < 3; d++)
        pr.t[d * pr.stride + i] = kind == RP_MOVE_ROTATION ? 0.0 : philox_uniform(rng, -ts, ts);
    double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (kind != RP_MOVE_TRANSLATION) {
        V3 axis;
        do {
            axis.x = philox_uniform(rng, -1, 1);
            axis.y = philox_uniform(rng, -1, 1);
            axis.z = philox_uniform(rng, -1, 1);
        } while (dot(axis, axis) > 1 || dot(axis, axis) == 0);
        rodrigues(axis, philox_uniform(rng, -rs, rs), R);
    }
    for (int c = 0; c < 9; c++) pr.R[c * pr.stride + i] = R[c];
    if (pr.u) pr.u[i] = philox_unit(rng);  // Simulation.cpp:483, drawn after the move as there
    if (pr.type) pr.type[i] = (uint8_t)type;
    if (pr.acc) pr.acc[i] = 0;
}

hipError_t launch_apply_moves(hipStream_t s, const double *dims, const double *inv, size_t n, const double *pos,
                              const double *rot, const double *t, const double *axis, const double *angle,
                              double *outPos, double *outR, double *outRot, double *outSc) {
    if (n == 0) return hipSuccess;
    M3 d, v;
    for (int c = 0; c < 9; c++) {
        d.m[c] = dims[c];
        v.m[c] = inv[c];
    }
    hipLaunchKernelGGL(k_apply_moves, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, v, n, pos, rot, t, axis,
                       angle, outPos, outR, outRot, outSc);
    return hipGetLastError();
}

hipError_t launch_proposals(hipStream_t s, const Grid &g, const Proposals &pr, int moves, const int *cellOf, int n,
                            const MoveSet &ms, uint64_t seed, uint32_t cycle, Timed t, FoldSlots fold) {
    const size_t total = moves > 0 ? (size_t)g.ncell * moves : (size_t)n * ms.n;
    if (!fold.slots) fold.cells = 0;
    const size_t threads = std::max(total, (size_t)fold.cells);
    if (threads == 0) return hipSuccess;
    hipExtLaunchKernelGGL(k_proposals, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, t.start, t.stop, 0,
                          g, pr, moves, cellOf, n, ms, seed, cycle, fold);
    return hipGetLastError();
}

// Per-sampler attempts and acceptances of one cycle's proposals, summed per workgroup in LDS.
__global__ __launch_bounds__(256) void k_type_counts(Proposals pr, size_t total, unsigned long long *counters) {
    __shared__ unsigned s[2 * RP_MAX_MOVE_SAMPLERS];
    if (threadIdx.x < 2 * RP_MAX_MOVE_SAMPLERS) s[threadIdx.x] = 0;
    __syncthreads();
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) {
        const int t = pr.type[i];
        if (t < RP_MAX_MOVE_SAMPLERS) {
            atomicAdd(&s[t], 1u);
            if (pr.acc[i]) atomicAdd(&s[RP_MAX_MOVE_SAMPLERS + t], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < RP_MAX_MOVE_SAMPLERS) {
        if (s[threadIdx.x]) atomicAdd(&counters[CNT_TYPE_ATT + threadIdx.x], (unsigned long long)s[threadIdx.x]);
        const unsigned a = s[RP_MAX_MOVE_SAMPLERS + threadIdx.x];
        if (a) atomicAdd(&counters[CNT_TYPE_ACC + threadIdx.x], (unsigned long long)a);
    }
}

hipError_t launch_type_counts(hipStream_t s, const Proposals &pr, size_t total, unsigned long long *counters) {
    if (total == 0 || !pr.type || !pr.acc) return hipSuccess;
    hipLaunchKernelGGL(k_type_counts, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, pr, total, counters);
    return hipGetLastError();
}

// LDS of one sweep workgroup: the stencil (cap entries), the ambiguous-candidate queues (cap + 64*waves ints),
// own orientations, one batch of proposals and the translation table.
// The guarded FP32 scan adds the FP32 stencil positions (3 x cap floats) and the FP32 translation table (81 floats).
size_t sweep_lds_bytes(int kind, int cap, int waves, int g32) {
    const int queueInts = max(cap + WAVE * waves, (waves - 1) * GROUPED_QUEUE);
    return stencil_lds_bytes(kind, cap) + sizeof(int) * queueInts +
           sizeof(double) * (9 * OWN_CAP + 12 * MOVE_CHUNK + 81 + SWEEP_AUX) + sizeof(int) * (MOVE_CHUNK + SWEEP_AUX_I) +
           (g32 ? sizeof(float) * (3 * (size_t)cap + 81) : 0);
}

__device__ __forceinline__ void flush_counters(const SweepParams &P, int cell, unsigned long long accepted, int moves,
                                               unsigned long long leftCell, unsigned long long tests, bool staged) {
    if (threadIdx.x == 0) {
        unsigned long long *slot = P.slots + (size_t)cell * SLOT_WORDS;
        slot[SLOT_ACCEPTED] = accepted;
        slot[SLOT_ATTEMPTED] = (unsigned long long)moves;
        slot[SLOT_LEFT_CELL] = leftCell;
        slot[SLOT_TESTS] = tests;
        if (!staged) atomicAdd(&P.counters[CNT_UNSTAGED_BLOCKS], 1ull);
    }
}


// Copies proposals [first, first + cn) of this cell into LDS.
__device__ __forceinline__ void load_proposals(const SweepParams &P, int pbase, int first, int cn, double *pt,
                                               double *pR, int *pk) {
    const int tid = threadIdx.x;
    if (tid < cn) {
        const int i = pbase + first + tid;
        pk[tid] = P.pr.k[i];
        for (int d = 0; d < 3; d++) pt[d * MOVE_CHUNK + tid] = P.pr.t[d * P.pr.stride + i];
        for (int c = 0; c < 9; c++) pR[c * MOVE_CHUNK + tid] = P.pr.R[c * P.pr.stride + i];
    }
}

// One block-wide vote per move with a single barrier: a wavefront that saw a hit stores the move's tag in the
// move's slot (two slots alternate; every move passes exactly one barrier, so a slot is rewritten only after
// every lane has read it).
__device__ __forceinline__ bool block_vote(bool hit, int move, volatile int *slots) {
    if (__any(hit) && (threadIdx.x & (WAVE - 1)) == 0) slots[move & 1] = move + 1;
    __syncthreads();
    return slots[move & 1] == move + 1;
}

// Staged sweep of one active cell.  TRANSLATED: the cell is not in the last row of any axis, so no own particle
// has a wrap bit set and every candidate was staged as pos + T (the reference's pos2bc, rounded once), so the
// trial pose is compared against the staged positions directly.  Each trial move:
//   pass A: every lane loads its candidates, computes the centre distance^2 exactly as
//           SpherocylinderTraits::overlapBetween does, marks definite overlaps (d^2 < 4r^2) and appends the
//           ambiguous ones (4r^2 <= d^2 < (2r+L)^2) to its wavefront's LDS queue;
//   pass B: the wavefront runs the exact segment-distance test on its queue (skipped once it saw a definite hit);
//   one workgroup vote decides the move.  An accepted move is written into LDS by every lane (same values), so
//   no second barrier is needed before the next move reads it.
template <int KIND, int WAVES, bool TRANSLATED>
__device__ __forceinline__ void sweep_cell_staged(const SweepParams &P, const Stencil &S, double *oo, double *pt, double *pR,
                                  int *pk, int *queue, const double *sT, volatile int *vote, int cell, int own0,
                                  int nOwn, int total, unsigned long long *clk) {
    constexpr int NT = WAVE * WAVES;
    const Grid &g = P.g;
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int wave = tid / WAVE;
    const int slots = (total + NT - 1) / NT;
    int *myQueue = queue + wave * WAVE * slots;
    unsigned long long accepted = 0, leftCell = 0, tests = 0;
    int move = 0;
    const int moves = P.moves > 0 ? P.moves : nOwn * P.mult;
    const int pbase = P.moves > 0 ? cell * P.moves : own0 * P.mult;

    for (int chunk = 0; chunk < moves; chunk += MOVE_CHUNK) {
        const int cn = min(MOVE_CHUNK, moves - chunk);
        if (P.fixMode) {
            if (tid < cn) pk[tid] = chunk + tid;
        } else {
            load_proposals(P, pbase, chunk, cn, pt, pR, pk);
        }
        __syncthreads();

        for (int mi = 0; mi < cn; mi++, move++) {
            const int kIdx = pk[mi];
            V3 pos = v3(S.x[kIdx], S.y[kIdx], S.z[kIdx]);
            V3 npos = pos;
            if (!P.fixMode) {
                V3 t = v3(pt[mi], pt[MOVE_CHUNK + mi], pt[2 * MOVE_CHUNK + mi]);
                npos = translate_wrapped(g.dims, g.inv, pos, t);
            }
            int nc[3];
            unsigned nw;
            bin_position(g, npos, nc, nw);
            const bool inside = cell_index(g, nc[0], nc[1], nc[2]) == cell;
            double NO[9];
            {
                double O[9], R[9];
                for (int c = 0; c < 9; c++) {
                    O[c] = oo[c * OWN_CAP + kIdx];
                    R[c] = P.fixMode ? 0.0 : pR[c * MOVE_CHUNK + mi];
                }
                if (P.fixMode) {
                    for (int c = 0; c < 9; c++) NO[c] = O[c];
                    fix_rotation(NO);  // Packing::tryOrientationFix (Packing.cpp:1218-1240)
                } else {
                    mulm(R, O, NO);  // Shape::rotate (Shape.cpp:14-16)
                }
            }
            const V3 nax = v3(NO[2], NO[5], NO[8]);
            const int wx = nw & 1, wy = (nw >> 1) & 1, wz = (nw >> 2) & 1;
            bool hit = !inside;
            if (inside) {
                SCFirst f = make_first<KIND>(P.k, npos, nax);
                // pass A, SLOT_GROUP candidates per lane loaded before any is tested
                bool definite = false;
                int queued = 0;
                for (int s0 = 0; s0 < slots; s0 += SLOT_GROUP) {
                    V3 p2[SLOT_GROUP];
                    int jj[SLOT_GROUP];
#pragma unroll
                    for (int s = 0; s < SLOT_GROUP; s++) {
                        jj[s] = (s0 + s) * NT + tid;
                        const int j = jj[s] < total ? jj[s] : 0;
                        p2[s] = v3(S.x[j], S.y[j], S.z[j]);
                        if (!TRANSLATED) p2[s] = add(p2[s], image_translation(sT, S.code[j], wx, wy, wz));
                    }
#pragma unroll
                    for (int s = 0; s < SLOT_GROUP; s++) {
                        const int j = jj[s];
                        bool amb = false;
                        if (j < total && j != kIdx) {
                            V3 d = sub(p2[s], npos);
                            double d2 = dot(d, d);
                            if (d2 < P.k.contact2) definite = true;
                            else if (KIND == 1 && d2 < P.k.range2) amb = true;
                        }
                        if (KIND == 1 && s0 + s < slots) {
                            unsigned long long m = __ballot(amb);
                            if (amb) myQueue[queued + __popcll(m & ((1ull << lane) - 1))] = j;
                            queued += __popcll(m);
                        }
                    }
                }
                tests += (unsigned long long)total;
                __builtin_amdgcn_wave_barrier();  // the queue was written by other lanes of this wavefront
                hit = definite;
                // pass B
                if (KIND == 1 && !__any(definite)) {
                    for (int q = lane; q < queued; q += WAVE) {
                        const int j = myQueue[q];
                        V3 p2 = v3(S.x[j], S.y[j], S.z[j]);
                        if (!TRANSLATED) p2 = add(p2, image_translation(sT, S.code[j], wx, wy, wz));
                        V3 a2 = v3(S.ax[j], S.ay[j], S.az[j]);
                        V3 c2a = add(p2, scal(a2, -P.k.halfLen));
                        V3 c2b = add(p2, scal(a2, P.k.halfLen));
                        if (segment_distance2(f.capA, f.u, f.a, c2a, c2b) < P.k.contact2) hit = true;
                    }
                }
            } else {
                leftCell++;
            }
            if (block_vote(hit, move, vote)) continue;

            // Packing::acceptMove: the cell belongs to this workgroup for the whole colour pass
            accepted++;
            if (WAVES < 8 && P.pr.acc && !P.fixMode && tid == 0) P.pr.acc[pbase + chunk + mi] = 1;
            S.x[kIdx] = npos.x;
            S.y[kIdx] = npos.y;
            S.z[kIdx] = npos.z;
            if (KIND == 1) {
                S.ax[kIdx] = nax.x;
                S.ay[kIdx] = nax.y;
                S.az[kIdx] = nax.z;
            }
            S.code[kIdx] = (wx + 1) | ((wy + 1) << 2) | ((wz + 1) << 4);
            for (int c = 0; c < 9; c++) oo[c * OWN_CAP + kIdx] = NO[c];
        }
        __syncthreads();
    }
    if (clk) clk[3] = wall_clock64();
    for (int a = tid; a < nOwn; a += NT) {
        const int gi = own0 + a;
        P.p.x[gi] = S.x[a];
        P.p.y[gi] = S.y[a];
        P.p.z[gi] = S.z[a];
        for (int c = 0; c < 9; c++) P.p.o[c][gi] = oo[c * OWN_CAP + a];
        int code = S.code[a];
        P.p.wbits[gi] = (uint8_t)(((code & 3) - 1) | ((((code >> 2) & 3) - 1) << 1) | ((((code >> 4) & 3) - 1) << 2));
    }
    flush_counters(P, cell, accepted, moves, leftCell, tests, true);
}

// Pipelined sweep of one staged cell (two or more wavefronts).  Wavefront 0 is the planner: while the other
// wavefronts (testers) test trial move m against the stencil, it applies proposal m+1 to the current pose --
// PBC wrap, cell check, R * O, cap geometry -- and publishes the plan in LDS; the move vote is the only
// workgroup barrier per move.  If move m is accepted and move m+1 picks the same particle, the plan was made
// from the old pose and is recomputed after the vote.
constexpr int PLAN_D = 24;  // doubles per plan: npos 3, nax 3, capA 3, u 3, a 1, NO 9, spare 2

template <int KIND>
__device__ __forceinline__ void plan_move(const SweepParams &P, const Stencil &S, const double *oo, const double *pt,
                                          const double *pR, const int *pk, int mi, int cell, double *plan,
                                          int *planI) {
    const Grid &g = P.g;
    const int kIdx = pk[mi];
    V3 pos = v3(S.x[kIdx], S.y[kIdx], S.z[kIdx]);
    V3 t = v3(pt[mi], pt[MOVE_CHUNK + mi], pt[2 * MOVE_CHUNK + mi]);
    V3 npos = translate_wrapped(g.dims, g.inv, pos, t);
    int nc[3];
    unsigned nw;
    bin_position(g, npos, nc, nw);
    const int inside = cell_index(g, nc[0], nc[1], nc[2]) == cell;
    double O[9], R[9], NO[9];
    for (int c = 0; c < 9; c++) {
        O[c] = oo[c * OWN_CAP + kIdx];
        R[c] = pR[c * MOVE_CHUNK + mi];
    }
    mulm(R, O, NO);  // Shape::rotate (Shape.cpp:14-16)
    const V3 nax = v3(NO[2], NO[5], NO[8]);
    SCFirst f = make_first<KIND>(P.k, npos, nax);
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        plan[0] = npos.x; plan[1] = npos.y; plan[2] = npos.z;
        plan[3] = nax.x; plan[4] = nax.y; plan[5] = nax.z;
        if (KIND == 1) {
            plan[6] = f.capA.x; plan[7] = f.capA.y; plan[8] = f.capA.z;
            plan[9] = f.u.x; plan[10] = f.u.y; plan[11] = f.u.z;
            plan[12] = f.a;
        }
        for (int c = 0; c < 9; c++) plan[13 + c] = NO[c];
        planI[0] = kIdx;
        planI[1] = (int)nw;
        planI[2] = inside;
    }
}

template <int KIND, int WAVES, bool TRANSLATED>
__device__ __forceinline__ void sweep_cell_pipelined(const SweepParams &P, const Stencil &S, double *oo, double *pt, double *pR,
                                     int *pk, int *queue, const double *sT, volatile int *vote, double *plans,
                                     int *plansI, int own0, int nOwn, int total, int cell,
                                     unsigned long long *clk) {
    constexpr int NTT = WAVE * (WAVES - 1);  // tester lanes
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int wave = tid / WAVE;
    const bool planner = wave == 0;
    const int ttid = tid - WAVE;  // tester index
    const int slots = (total + NTT - 1) / NTT;
    int *myQueue = queue + (wave > 0 ? (wave - 1) : 0) * WAVE * slots;
    unsigned long long accepted = 0, leftCell = 0, tests = 0;
    int move = 0;
    const int moves = P.moves > 0 ? P.moves : nOwn * P.mult;
    const int pbase = P.moves > 0 ? cell * P.moves : own0 * P.mult;

    for (int chunk = 0; chunk < moves; chunk += MOVE_CHUNK) {
        const int cn = min(MOVE_CHUNK, moves - chunk);
        load_proposals(P, pbase, chunk, cn, pt, pR, pk);
        __syncthreads();
        if (planner) plan_move<KIND>(P, S, oo, pt, pR, pk, 0, cell, plans, plansI);
        __syncthreads();

        for (int mi = 0; mi < cn; mi++, move++) {
            double *plan = plans + (mi & 1) * PLAN_D;
            int *planI = plansI + (mi & 1) * 4;
            // the pose is read before the vote: once a wavefront has passed it, the planner may overwrite this
            // plan buffer with the plan of move mi + 2
            const int kIdx = planI[0];
            const int nw = planI[1];
            const bool inside = planI[2] != 0;
            const V3 npos = v3(plan[0], plan[1], plan[2]);
            const V3 nax = v3(plan[3], plan[4], plan[5]);
            bool hit = !inside;
            if (planner) {
                if (mi + 1 < cn)
                    plan_move<KIND>(P, S, oo, pt, pR, pk, mi + 1, cell, plans + ((mi + 1) & 1) * PLAN_D,
                                    plansI + ((mi + 1) & 1) * 4);
            } else if (inside) {
                const int wx = nw & 1, wy = (nw >> 1) & 1, wz = (nw >> 2) & 1;
                bool definite = false;
                int queued = 0;
                for (int s0 = 0; s0 < slots; s0 += SLOT_GROUP) {
                    V3 p2[SLOT_GROUP];
                    int jj[SLOT_GROUP];
#pragma unroll
                    for (int s = 0; s < SLOT_GROUP; s++) {
                        jj[s] = (s0 + s) * NTT + ttid;
                        const int j = jj[s] < total ? jj[s] : 0;
                        if (s0 + s < slots) {
                            p2[s] = v3(S.x[j], S.y[j], S.z[j]);
                            if (!TRANSLATED) p2[s] = add(p2[s], image_translation(sT, S.code[j], wx, wy, wz));
                        }
                    }
#pragma unroll
                    for (int s = 0; s < SLOT_GROUP; s++) {
                        if (s0 + s >= slots) continue;
                        const int j = jj[s];
                        bool amb = false;
                        if (j < total && j != kIdx) {
                            V3 d = sub(p2[s], npos);
                            double d2 = dot(d, d);
                            if (d2 < P.k.contact2) definite = true;
                            else if (KIND == 1 && d2 < P.k.range2) amb = true;
                        }
                        if (KIND == 1) {
                            unsigned long long m = __ballot(amb);
                            if (amb) myQueue[queued + __popcll(m & ((1ull << lane) - 1))] = j;
                            queued += __popcll(m);
                        }
                    }
                }
                __builtin_amdgcn_wave_barrier();  // the queue was written by other lanes of this wavefront
                hit = definite;
                if (KIND == 1 && !__any(definite)) {
                    SCFirst f;
                    f.capA = v3(plan[6], plan[7], plan[8]);
                    f.u = v3(plan[9], plan[10], plan[11]);
                    f.a = plan[12];
                    for (int q = lane; q < queued; q += WAVE) {
                        const int j = myQueue[q];
                        V3 p2 = v3(S.x[j], S.y[j], S.z[j]);
                        if (!TRANSLATED) p2 = add(p2, image_translation(sT, S.code[j], wx, wy, wz));
                        V3 a2 = v3(S.ax[j], S.ay[j], S.az[j]);
                        V3 c2a = add(p2, scal(a2, -P.k.halfLen));
                        V3 c2b = add(p2, scal(a2, P.k.halfLen));
                        if (segment_distance2(f.capA, f.u, f.a, c2a, c2b) < P.k.contact2) hit = true;
                    }
                }
            }
            if (!inside) leftCell++;
            else tests += (unsigned long long)total;
            if (block_vote(hit, move, vote)) continue;

            // Packing::acceptMove: every lane writes the same values, so the next move sees them without a barrier
            accepted++;
            if (WAVES < 8 && P.pr.acc && tid == 0) P.pr.acc[pbase + chunk + mi] = 1;
            S.x[kIdx] = npos.x;
            S.y[kIdx] = npos.y;
            S.z[kIdx] = npos.z;
            if (KIND == 1) {
                S.ax[kIdx] = nax.x;
                S.ay[kIdx] = nax.y;
                S.az[kIdx] = nax.z;
            }
            S.code[kIdx] = ((nw & 1) + 1) | ((((nw >> 1) & 1) + 1) << 2) | ((((nw >> 2) & 1) + 1) << 4);
            if (planner) {
                for (int c = 0; c < 9; c++) oo[c * OWN_CAP + kIdx] = plan[13 + c];
            }
            // every lane of the planner has read this plan buffer before its lane 0 rewrites it with the plan of
            // move mi + 2 (a scheduling fence; no instruction)
            __builtin_amdgcn_wave_barrier();
            if (mi + 1 < cn && pk[mi + 1] == kIdx) {
                // the next plan used the pose this move just replaced
                __syncthreads();
                if (planner)
                    plan_move<KIND>(P, S, oo, pt, pR, pk, mi + 1, cell, plans + ((mi + 1) & 1) * PLAN_D,
                                    plansI + ((mi + 1) & 1) * 4);
                __syncthreads();
            }
        }
        __syncthreads();
    }
    if (clk) clk[3] = wall_clock64();
    for (int a = tid; a < nOwn; a += WAVE * WAVES) {
        const int gi = own0 + a;
        P.p.x[gi] = S.x[a];
        P.p.y[gi] = S.y[a];
        P.p.z[gi] = S.z[a];
        for (int c = 0; c < 9; c++) P.p.o[c][gi] = oo[c * OWN_CAP + a];
        int code = S.code[a];
        P.p.wbits[gi] = (uint8_t)(((code & 3) - 1) | ((((code >> 2) & 3) - 1) << 1) | ((((code >> 4) & 3) - 1) << 2));
    }
    flush_counters(P, cell, accepted, moves, leftCell, tests, true);
}

// Sweep of one active cell whose stencil does not fit the LDS area: candidates are read from the cell-sorted
// arrays in HBM, accepted moves are written there at once.
// Polysphere candidate test: particle j at p2 (orientation in p) under translation T against the trial beads of the
// moved particle (sTrial, centred at npos).  Candidates whose centre is at least the prefilter distance away
// (k.range2 = (total range (1 + 1e-9))^2) cannot reach any bead; otherwise every bead pair is tested as
// PolysphereTraits::HardInteraction::overlapBetween with the candidate's beads at p2 + O_j c_b.
// Whether any bead pair of the moved particle (beads sTrial, centre npos) and candidate j (p2 + O_j c_b under
// translation T) overlaps.  Two exact filters skip pairs that cannot: each particle's beads lie within k.bound of
// its centre, so a bead of either particle at least bound + its own radius from the other particle's centre is
// at least the sum of radii from every bead of the other.  The skip needs that distance squared to exceed
// (bound + r)^2 by a relative 1e-9, far above the rounding of positions and of the bead test's own distance,
// so the decision equals the full 11 x 11 test's.
__device__ __forceinline__ bool poly_hit(const PairConsts &k, const double *sBeads, const double *sTrial,
                                         const V3 &npos, const V3 &p2, const Particles &p, int gj, const V3 &T) {
    const V3 q2 = add(p2, T);
    const V3 d = sub(q2, npos);
    if (!(dot(d, d) < k.range2)) return false;
    // moved beads within reach of the candidate's bounding sphere (nc <= 16)
    unsigned near = 0;
    for (int a = 0; a < k.nc; a++) {
        const V3 e = sub(v3(sTrial[3 * a], sTrial[3 * a + 1], sTrial[3 * a + 2]), q2);
        const double t = k.bound + sBeads[4 * a + 3];
        if (!(dot(e, e) > (t * t) * (1 + 1e-9))) near |= 1u << a;
    }
    if (!near) return false;
    double O[9];
    for (int e = 0; e < 9; e++) O[e] = p.o[e][gj];
    for (int b = 0; b < k.nc; b++) {
        const V3 c2 = add(p2, mulv(O, v3(sBeads[4 * b], sBeads[4 * b + 1], sBeads[4 * b + 2])));
        const V3 f = sub(add(c2, T), npos);
        const double t = k.bound + sBeads[4 * b + 3];
        if (dot(f, f) > (t * t) * (1 + 1e-9)) continue;  // candidate bead out of reach of the moved particle
        for (unsigned m = near; m; m &= m - 1) {
            const int a = __builtin_ctz(m);
            const V3 c1 = v3(sTrial[3 * a], sTrial[3 * a + 1], sTrial[3 * a + 2]);
            if (bead_overlap(sBeads[4 * a + 3], sBeads[4 * b + 3], c1, c2, T)) return true;
        }
    }
    return false;
}

// Overlapping bead pairs of particle j (p2 + O_j c_b under translation T) with beads sBeadPos (the moved
// particle's beads, centred at pos); the count of Packing::countParticleOverlaps over interaction centres
// (Packing.cpp:381-409, 619-647).  Candidates at least the prefilter distance away contribute none.
__device__ __forceinline__ int poly_count(const PairConsts &k, const double *sBeads, const double *sBeadPos,
                                          const V3 &pos, const V3 &p2, const Particles &p, int gj, const V3 &T) {
    const V3 d = sub(add(p2, T), pos);
    if (!(dot(d, d) < k.range2)) return 0;
    double O[9];
    for (int e = 0; e < 9; e++) O[e] = p.o[e][gj];
    int count = 0;
    for (int b = 0; b < k.nc; b++) {
        const V3 c2 = add(p2, mulv(O, v3(sBeads[4 * b], sBeads[4 * b + 1], sBeads[4 * b + 2])));
        for (int a = 0; a < k.nc; a++) {
            const V3 c1 = v3(sBeadPos[3 * a], sBeadPos[3 * a + 1], sBeadPos[3 * a + 2]);
            count += bead_overlap(sBeads[4 * a + 3], sBeads[4 * b + 3], c1, c2, T) ? 1 : 0;
        }
    }
    return count;
}

template <int KIND, int WAVES>
__device__ __forceinline__ void sweep_cell_unstaged(const SweepParams &P, const Stencil &S, double *pt, double *pR, int *pk,
                                    const double *sT, int cell, int own0, int nOwn, int total) {
    constexpr int NT = WAVE * WAVES;
    const Grid &g = P.g;
    const int tid = threadIdx.x;
    unsigned long long accepted = 0, leftCell = 0, tests = 0;
    const int moves = P.moves > 0 ? P.moves : nOwn * P.mult;
    const int pbase = P.moves > 0 ? cell * P.moves : own0 * P.mult;
    // polyspheres: the bead table, and the trial pose's bead centres of the current move
    __shared__ double sBeads[KIND == 2 ? 4 * RP_MAX_CENTRES : 1], sTrial[KIND == 2 ? 3 * RP_MAX_CENTRES : 1];
    if (KIND == 2) {
        for (int e = tid; e < 4 * P.k.nc; e += NT) sBeads[e] = P.k.beads[e];
        __syncthreads();
    }
    for (int chunk = 0; chunk < moves; chunk += MOVE_CHUNK) {
        const int cn = min(MOVE_CHUNK, moves - chunk);
        if (P.fixMode) {
            if (tid < cn) pk[tid] = chunk + tid;
        } else {
            load_proposals(P, pbase, chunk, cn, pt, pR, pk);
        }
        __syncthreads();
        for (int mi = 0; mi < cn; mi++) {
            const int kIdx = pk[mi];
            const int gi = own0 + kIdx;
            V3 pos = load3(P.p.x, P.p.y, P.p.z, gi);
            V3 npos = pos;
            if (!P.fixMode) {
                V3 t = v3(pt[mi], pt[MOVE_CHUNK + mi], pt[2 * MOVE_CHUNK + mi]);
                npos = translate_wrapped(g.dims, g.inv, pos, t);
            }
            int nc[3];
            unsigned nw;
            bin_position(g, npos, nc, nw);
            if (cell_index(g, nc[0], nc[1], nc[2]) != cell) {
                leftCell++;
                continue;
            }
            double O[9], R[9], NO[9];
            for (int c = 0; c < 9; c++) {
                O[c] = P.p.o[c][gi];
                R[c] = P.fixMode ? 0.0 : pR[c * MOVE_CHUNK + mi];
            }
            if (P.fixMode) {
                for (int c = 0; c < 9; c++) NO[c] = O[c];
                fix_rotation(NO);
            } else {
                mulm(R, O, NO);
            }
            V3 nax = v3(NO[2], NO[5], NO[8]);
            SCFirst f = make_first<KIND>(P.k, npos, nax);
            const int wx = nw & 1, wy = (nw >> 1) & 1, wz = (nw >> 2) & 1;
            if (KIND == 2) {
                // trial bead centres npos + NO * c (PolysphereTraits.cpp:93-95)
                if (tid < P.k.nc) {
                    const V3 c1 = add(npos, mulv(NO, v3(sBeads[4 * tid], sBeads[4 * tid + 1], sBeads[4 * tid + 2])));
                    sTrial[3 * tid] = c1.x;
                    sTrial[3 * tid + 1] = c1.y;
                    sTrial[3 * tid + 2] = c1.z;
                }
                __syncthreads();
            }
            bool hit = false;
            for (int base = 0; base < total; base += NT) {
                const int j = base + tid;
                bool h = false;
                if (j < total) {
                    // the moved particle is skipped under every image: with one cell across an axis its own cell
                    // recurs in the stencil (Packing.cpp:383-384 skips the particle's own index in every cell)
                    const int q = stencil_cell(S.start, S.ncells, j);
                    const int gj = S.base[q] + (j - S.start[q]);
                    if (gj != gi) {
                        const V3 p2 = load3(P.p.x, P.p.y, P.p.z, gj);
                        const V3 T = image_translation(sT, image_code(S.ecode[q], P.p.wbits[gj]), wx, wy, wz);
                        if (KIND == 2) {
                            h = poly_hit(P.k, sBeads, sTrial, npos, p2, P.p, gj, T);
                        } else {
                            const V3 a2 = KIND == 1 ? v3(P.p.o[2][gj], P.p.o[5][gj], P.p.o[8][gj]) : v3(0, 0, 0);
                            h = pair_overlap<KIND>(P.k, f, p2, a2, T);
                        }
                    }
                }
                tests += (unsigned long long)min(NT, total - base);
                bool any;
                if (WAVES == 1) any = __any(h);
                else any = __syncthreads_or(h) != 0;
                if (any) {
                    hit = true;
                    break;
                }
            }
            if (hit) continue;
            accepted++;
            if (P.pr.acc && !P.fixMode && tid == 0) P.pr.acc[pbase + chunk + mi] = 1;
            if (tid == 0) {
                P.p.x[gi] = npos.x;
                P.p.y[gi] = npos.y;
                P.p.z[gi] = npos.z;
                for (int c = 0; c < 9; c++) P.p.o[c][gi] = NO[c];
                P.p.wbits[gi] = (uint8_t)nw;
            }
            __syncthreads();
        }
        __syncthreads();
    }
    flush_counters(P, cell, accepted, moves, leftCell, tests, false);
}

// Grouped sweep of one staged cell (two or more wavefronts).  Up to GROUP consecutive trial moves on distinct
// particles are tested against the stencil at once, each candidate loaded once for all of them, and then decided in
// their order exactly as the sequential sweep decides them: move k is rejected if it overlaps a particle outside the
// group, or a group particle i in the pose that particle has when move k runs -- the old pose for i > k or for a
// rejected move i < k (found by the stencil scan), the trial pose for an accepted move i < k (found by the planner's
// pair tests between trial poses).  Same decisions, same trajectory as one move at a time.  Wavefront 0 plans the
// next group (one lane per move) while the others scan; one workgroup barrier per group.  Plans and hit words are
// triple-buffered: a buffer is rewritten two groups after its use, when every wavefront has passed the barrier that
// follows its last read.
constexpr int GROUP = 4;
constexpr int GQ = GROUPED_QUEUE;
constexpr int GPLAN_I = 8;       // ints per group plan: kIdx[4], size, trial-pair hits, packed (nw | inside << 3)
constexpr int GBUF = 3;          // plan / hit-word buffers
constexpr int PK_ACCEPTED = 1 << 30;  // set on a decided move's particle index in LDS once it is accepted
static_assert(SWEEP_AUX >= GBUF * GROUP * PLAN_D && SWEEP_AUX_I >= GBUF * GPLAN_I + GBUF, "grouped layout");

template <int KIND>
__device__ __forceinline__ bool hard_pair_hit(const PairConsts &k, const SCFirst &f, const V3 &npos, const V3 &p2,
                                              const V3 &a2) {
    V3 d = sub(p2, npos);
    double d2 = dot(d, d);
    if (d2 < k.contact2) return true;
    if (KIND == 0 || !(d2 < k.range2)) return false;
    V3 c2a = add(p2, scal(a2, -k.halfLen));
    V3 c2b = add(p2, scal(a2, k.halfLen));
    return segment_distance2(f.capA, f.u, f.a, c2a, c2b) < k.contact2;
}

// Size of the group starting at move m0 of the chunk: the longest prefix of at most GROUP moves on distinct
// particles.
__device__ __forceinline__ int group_size(const int *pk, int m0, int cn) {
    int gs = 1;
    while (gs < GROUP && m0 + gs < cn) {
        const int kk = pk[m0 + gs];
        bool dup = false;
        for (int i = 0; i < gs; i++) dup |= pk[m0 + i] == kk;
        if (dup) break;
        gs++;
    }
    return gs;
}

__device__ __forceinline__ SCFirst plan_first(const double *q) {
    SCFirst f;
    f.pos = v3(q[0], q[1], q[2]);
    f.capA = v3(q[6], q[7], q[8]);
    f.u = v3(q[9], q[10], q[11]);
    f.a = q[12];
    return f;
}

__device__ __forceinline__ int wrap_code(unsigned nw) {
    return (int)((nw & 1) + 1) | ((int)(((nw >> 1) & 1) + 1) << 2) | ((int)(((nw >> 2) & 1) + 1) << 4);
}

// Planner wavefront: lane t < gs plans move m0 + t into plan slot t; then lanes 0-5 test the trial poses against
// each other (bit 4k + i: move k's trial pose overlaps move i's trial pose, i < k).
// With G32 the plan also carries the trial position relative to o in FP32 (three floats in doubles 22-23).
template <int KIND, bool TRANSLATED, bool G32 = false>
__device__ __forceinline__ void plan_group(const SweepParams &P, const Stencil &S, const double *oo, const double *pt,
                                           const double *pR, const int *pk, const double *sT, int m0, int gs,
                                           int cell, double *plan, int *planI, V3 o = V3{}) {
    const Grid &g = P.g;
    const int lane = threadIdx.x & (WAVE - 1);
    unsigned byte = 0;
    if (lane < gs) {
        const int mi = m0 + lane;
        const int kIdx = pk[mi];
        V3 pos = v3(S.x[kIdx], S.y[kIdx], S.z[kIdx]);
        V3 t = v3(pt[mi], pt[MOVE_CHUNK + mi], pt[2 * MOVE_CHUNK + mi]);
        V3 npos = translate_wrapped(g.dims, g.inv, pos, t);
        int nc[3];
        unsigned nw;
        bin_position(g, npos, nc, nw);
        const unsigned inside = cell_index(g, nc[0], nc[1], nc[2]) == cell ? 1u : 0u;
        byte = nw | (inside << 3);
        double O[9], R[9], NO[9];
        for (int c = 0; c < 9; c++) {
            O[c] = oo[c * OWN_CAP + kIdx];
            R[c] = pR[c * MOVE_CHUNK + mi];
        }
        mulm(R, O, NO);  // Shape::rotate (Shape.cpp:14-16)
        const V3 nax = v3(NO[2], NO[5], NO[8]);
        SCFirst f = make_first<KIND>(P.k, npos, nax);
        double *q = plan + lane * PLAN_D;
        q[0] = npos.x; q[1] = npos.y; q[2] = npos.z;
        q[3] = nax.x; q[4] = nax.y; q[5] = nax.z;
        if (KIND == 1) {
            q[6] = f.capA.x; q[7] = f.capA.y; q[8] = f.capA.z;
            q[9] = f.u.x; q[10] = f.u.y; q[11] = f.u.z;
            q[12] = f.a;
        }
        for (int c = 0; c < 9; c++) q[13 + c] = NO[c];
        if (G32) {
            float *qf = reinterpret_cast<float *>(q + 22);
            qf[0] = (float)(npos.x - o.x);
            qf[1] = (float)(npos.y - o.y);
            qf[2] = (float)(npos.z - o.z);
        }
        planI[lane] = kIdx;
    }
    unsigned word = 0;
    for (int t = 0; t < gs; t++) word |= ((unsigned)__shfl((int)byte, t) & 15u) << (4 * t);
    __builtin_amdgcn_wave_barrier();  // the plans were written by other lanes of this wavefront
    unsigned hit = 0;
    if (lane < 6) {
        const int kk = lane < 1 ? 1 : (lane < 3 ? 2 : 3);
        const int ii = lane < 1 ? 0 : (lane < 3 ? lane - 1 : lane - 3);
        if (kk < gs) {
            const double *qk = plan + kk * PLAN_D, *qi = plan + ii * PLAN_D;
            SCFirst f;
            if (KIND == 1) f = plan_first(qk);
            else f.pos = v3(qk[0], qk[1], qk[2]);
            V3 p2 = v3(qi[0], qi[1], qi[2]);
            const V3 a2 = v3(qi[3], qi[4], qi[5]);
            if (!TRANSLATED) {
                const unsigned wk = (word >> (4 * kk)) & 7, wi = (word >> (4 * ii)) & 7;
                p2 = add(p2, image_translation(sT, wrap_code(wi), wk & 1, (wk >> 1) & 1, (wk >> 2) & 1));
            }
            if (hard_pair_hit<KIND>(P.k, f, f.pos, p2, a2)) hit = 1u << (4 * kk + ii);
        }
    }
    // lane l < 6 tests pair (kk, ii) = (1,0) (2,0) (2,1) (3,0) (3,1) (3,2): bits 4, 8, 9, 12, 13, 14
    const unsigned long long hm = __ballot(hit != 0);
    const unsigned pairs = (unsigned)(((hm & 1) << 4) | ((hm & 6) << 7) | ((hm & 56) << 9));
    if (lane == 0) {
        planI[4] = gs;
        planI[5] = (int)pairs;
        planI[6] = (int)word;
    }
}

// G32 = 1: the stencil scan classifies each (candidate, move) pair from the FP32 squared centre distance (positions
// relative to o); a pair whose FP32 value lies inside a guard band around contact^2 or range^2 (g32Lim) is decided
// by the FP64 distance, the same value the FP64 scan computes, so every decision equals the FP64 sweep's.  The
// bands are wider than the largest FP32 error the positions' magnitudes allow.  G32 = 2 also evaluates every pair in
// FP64 and records the largest FP32 error and any class that differs (none may).
template <int KIND, int WAVES, bool TRANSLATED, int G32 = 0>
__device__ __forceinline__ void sweep_cell_grouped(const SweepParams &P, const Stencil &S, double *oo, double *pt,
                                                   double *pR, int *pk, int *queue, const double *sT, double *plans,
                                                   int *plansI, int *hitWords, int own0, int nOwn, int total, int cell,
                                                   unsigned long long *clk, float *Sf = nullptr,
                                                   const float *sTf = nullptr, V3 o = V3{}) {
    constexpr int NTT = WAVE * (WAVES - 1);
    // a tester's queue: entries [0, GQ - WAVE), then one discard slot per lane for lanes with nothing to queue; a
    // slot adds at most GROUP * WAVE entries, so the queue is drained once it holds more than QFLUSH
    constexpr int QFLUSH = GQ - WAVE - GROUP * WAVE;
    static_assert(QFLUSH > 0, "grouped queue");
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int wave = tid / WAVE;
    const bool planner = wave == 0;
    const int ttid = tid - WAVE;
    int *myQueue = queue + (wave > 0 ? wave - 1 : 0) * GQ;
    int *discard = myQueue + (GQ - WAVE) + lane;
    unsigned long long accepted = 0, leftCell = 0, tests = 0;
    unsigned long long tBusy = 0, tDecide = 0, tTop = 0, tPassB = 0, nGroups = 0, nReplans = 0;  // phase timers
    unsigned g32Guard = 0, g32Mismatch = 0;  // guarded FP32 scan statistics of this lane
    unsigned long long g32Pairs = 0;         // pairs classified by this workgroup (one lane keeps the count)
    float g32Err = 0.f;
    const int moves = P.moves > 0 ? P.moves : nOwn * P.mult;
    const int pbase = P.moves > 0 ? cell * P.moves : own0 * P.mult;
    // issue priority: the planner and the testers with an extra slot bound each group's time
    if ((P.tuning & 1) && planner) __builtin_amdgcn_s_setprio(2);
    if ((P.tuning & 2) && !planner && (wave - 1) * WAVE < total - ((total - 1) / NTT) * NTT) __builtin_amdgcn_s_setprio(1);

    for (int chunk = 0; chunk < moves; chunk += MOVE_CHUNK) {
        const int cn = min(MOVE_CHUNK, moves - chunk);
        load_proposals(P, pbase, chunk, cn, pt, pR, pk);
        if (tid < GBUF) hitWords[tid] = 0;
        __syncthreads();
        // The chunk's groups, in registers of every wavefront: lane m holds the particle of move m and the length
        // of the run of moves on distinct particles that starts there (at most GROUP); the groups follow from
        // those runs (group_size), and lane g of `groups` holds group g as start | size << 8.
        const int pkv = lane < cn ? pk[lane] : -1;
        const int q1 = __shfl_down(pkv, 1), q2 = __shfl_down(pkv, 2), q3 = __shfl_down(pkv, 3);
        const bool d1 = lane + 1 < cn && q1 != pkv;
        const bool d2 = d1 && lane + 2 < cn && q2 != pkv && q2 != q1;
        const bool d3 = d2 && lane + 3 < cn && q3 != pkv && q3 != q1 && q3 != q2;
        const int runv = 1 + (int)d1 + (int)d2 + (int)d3;
        int groups = 0, ngroups = 0;
        for (int m = 0; m < cn; ngroups++) {
            const int r = __builtin_amdgcn_readlane(runv, m);
            groups = lane == ngroups ? m | (r << 8) : groups;
            m += r;
        }
        int gidx = 0;
        int gs = __builtin_amdgcn_readlane(groups, 0) >> 8;
        if (planner) plan_group<KIND, TRANSLATED, G32 != 0>(P, S, oo, pt, pR, pk, sT, 0, gs, cell, plans, plansI, o);
        __syncthreads();
        for (; gidx < ngroups; gidx++) {
            const unsigned long long tgTop = clk ? wall_clock64() : 0;
            const int buf = gidx % GBUF, nbuf = (gidx + 1) % GBUF;
            const double *plan = plans + buf * GROUP * PLAN_D;
            const int *planI = plansI + buf * GPLAN_I;
            // group sizes, plan words and particle indices are the same in every lane: kept in scalar registers
            const int m0 = __builtin_amdgcn_readlane(groups, gidx) & 255;
            gs = __builtin_amdgcn_readlane(groups, gidx) >> 8;
            const int m1 = m0 + gs;
            const int gsNext = gidx + 1 < ngroups ? __builtin_amdgcn_readlane(groups, gidx + 1) >> 8 : 0;
            int kIdx[GROUP];
            V3 npos[GROUP];
            float fm[GROUP][3];  // G32: trial positions relative to o
#pragma unroll
            for (int k = 0; k < GROUP; k++) {
                kIdx[k] = k < gs ? __builtin_amdgcn_readlane(pkv, m0 + k) : -1;
                if (G32) {
                    const float *qf = reinterpret_cast<const float *>(plan + k * PLAN_D + 22);
                    for (int c = 0; c < 3; c++) fm[k][c] = k < gs ? qf[c] : 0.f;
                } else {
                    npos[k] = k < gs ? v3(plan[k * PLAN_D], plan[k * PLAN_D + 1], plan[k * PLAN_D + 2]) : v3(0, 0, 0);
                }
            }
            const unsigned newHit = (unsigned)__builtin_amdgcn_readfirstlane(planI[5]);
            const unsigned word = (unsigned)__builtin_amdgcn_readfirstlane(planI[6]);  // nw | inside << 3 per move
            // bit k: group particle k is also moved by the next group, whose plan then needs its new pose
            unsigned nextUses = 0;
#pragma unroll
            for (int t = 0; t < GROUP; t++) {
                if (t < gsNext) {
                    const int kn = __builtin_amdgcn_readlane(pkv, m1 + t);
#pragma unroll
                    for (int k = 0; k < GROUP; k++) nextUses |= kn == kIdx[k] ? 1u << k : 0u;
                }
            }
            const unsigned long long tg0 = clk ? wall_clock64() : 0;
            if (planner) {
                if (gsNext)
                    plan_group<KIND, TRANSLATED, G32 != 0>(P, S, oo, pt, pR, pk, sT, m1, gsNext, cell,
                                                           plans + nbuf * GROUP * PLAN_D, plansI + nbuf * GPLAN_I, o);
                if (lane == 0) hitWords[nbuf] = 0;
            } else {
                // 5 bits per move: bits 0-3 overlap with group particle i's current pose, bit 4 with a particle
                // outside the group.  Each (candidate, move) pair is evaluated without branching on the lane: the
                // ballot of the ambiguous pairs is their compare mask, and lanes without one write a discard slot.
                unsigned hitAll = 0;
                int queued = 0;
                const int slots = (total + NTT - 1) / NTT;
                for (int s = 0; s < slots; s++) {
                    const int j = s * NTT + ttid;
                    const bool valid = j < total;
                    const int jj = valid ? j : 0;
                    const int code = TRANSLATED ? 0 : S.code[jj];
                    unsigned groupBit = 16;
#pragma unroll
                    for (int i = 0; i < GROUP; i++) groupBit = jj == kIdx[i] ? 1u << i : groupBit;
                    // the moves' distances first, as independent chains, then their hit bits and queue entries
                    double d2[GROUP];
                    float d2f[GROUP];
                    if (G32) {
                        const float fj[3] = {Sf[jj], Sf[P.cap + jj], Sf[2 * P.cap + jj]};
#pragma unroll
                        for (int k = 0; k < GROUP; k++) {
                            float q[3] = {fj[0], fj[1], fj[2]};
                            if (!TRANSLATED) {
                                const unsigned wk = (word >> (4 * k + 3)) & 1 ? (word >> (4 * k)) & 7 : 0u;
                                const int kx = (code & 3) - 1 - (int)(wk & 1), ky = ((code >> 2) & 3) - 1 - (int)((wk >> 1) & 1),
                                          kz = ((code >> 4) & 3) - 1 - (int)((wk >> 2) & 1);
                                const int t = 3 * ((kx + 1) * 9 + (ky + 1) * 3 + (kz + 1));
                                for (int c = 0; c < 3; c++) q[c] = q[c] + sTf[t + c];
                            }
                            const float dx = q[0] - fm[k][0], dy = q[1] - fm[k][1], dz = q[2] - fm[k][2];
                            d2f[k] = (dx * dx + dy * dy) + dz * dz;
                        }
                    } else {
                        const V3 p2 = v3(S.x[jj], S.y[jj], S.z[jj]);
#pragma unroll
                        for (int k = 0; k < GROUP; k++) {
                            V3 q2 = p2;
                            if (!TRANSLATED) {
                                // a move that leaves the cell is not tested; its wrap bits are zeroed so that the
                                // translation index stays inside the 27-entry table
                                const unsigned wk = (word >> (4 * k + 3)) & 1 ? (word >> (4 * k)) & 7 : 0u;
                                q2 = add(p2, image_translation(sT, code, wk & 1, (wk >> 1) & 1, (wk >> 2) & 1));
                            }
                            const V3 d = sub(q2, npos[k]);
                            d2[k] = dot(d, d);
                        }
                    }
#pragma unroll
                    for (int k = 0; k < GROUP; k++) {
                        if (k < gs && ((word >> (4 * k + 3)) & 1)) {  // moves leaving the cell are rejected
                            const bool other = valid && jj != kIdx[k];
                            bool close, within;
                            if (G32) {
                                const float *lim = P.g32Lim[TRANSLATED ? 0 : 1];
                                const float x = d2f[k];
                                close = x < lim[0];
                                within = x < lim[2];
                                const bool guard = (x >= lim[0] && x < lim[1]) || (KIND == 1 && x >= lim[2] && x < lim[3]);
                                if (guard || G32 == 2) {
                                    // the FP64 squared distance of the FP64 scan (rare outside verify mode)
                                    const double *qn = plan + k * PLAN_D;
                                    V3 q2 = v3(S.x[jj], S.y[jj], S.z[jj]);
                                    if (!TRANSLATED) {
                                        const unsigned wk = (word >> (4 * k)) & 7;
                                        q2 = add(q2, image_translation(sT, code, wk & 1, (wk >> 1) & 1, (wk >> 2) & 1));
                                    }
                                    const V3 d = sub(q2, v3(qn[0], qn[1], qn[2]));
                                    const double e2 = dot(d, d);
                                    const bool c64 = e2 < P.k.contact2, w64 = e2 < P.k.range2;
                                    if (G32 == 2 && other) {
                                        g32Err = fmaxf(g32Err, (float)fabs((double)x - e2));
                                        if (!guard && (c64 != close || (KIND == 1 && !c64 && w64 != within)))
                                            g32Mismatch++;
                                    }
                                    if (guard) {
                                        close = c64;
                                        within = w64;
                                        g32Guard += other ? 1u : 0u;
                                    }
                                }
                            } else {
                                close = d2[k] < P.k.contact2;
                                within = d2[k] < P.k.range2;
                            }
                            hitAll |= other && close ? groupBit << (5 * k) : 0u;
                            if (KIND == 1) {
                                const bool amb = other && !close && within;
                                const unsigned long long m = __ballot(amb);
                                const int at = queued + (int)__builtin_amdgcn_mbcnt_hi(
                                                            (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                                *(amb ? myQueue + at : discard) = (jj << 2) | k;
                                queued += __popcll(m);
                            }
                        }
                    }
                    if (KIND == 1 && queued > 0 && (queued > QFLUSH || s + 1 == slots)) {
                        // exact segment tests of the queued pairs (pass B)
                        const unsigned long long tb0 = clk ? wall_clock64() : 0;
                        __builtin_amdgcn_wave_barrier();  // the queue was written by other lanes of this wavefront
                        for (int q = lane; q < queued; q += WAVE) {
                            const int e = myQueue[q];
                            const int jq = e >> 2, k = e & 3;
                            V3 c2 = v3(S.x[jq], S.y[jq], S.z[jq]);
                            if (!TRANSLATED) {
                                const unsigned wk = (word >> (4 * k)) & 7;
                                c2 = add(c2, image_translation(sT, S.code[jq], wk & 1, (wk >> 1) & 1, (wk >> 2) & 1));
                            }
                            const V3 a2 = v3(S.ax[jq], S.ay[jq], S.az[jq]);
                            const SCFirst f = plan_first(plan + k * PLAN_D);
                            const V3 c2a = add(c2, scal(a2, -P.k.halfLen));
                            const V3 c2b = add(c2, scal(a2, P.k.halfLen));
                            if (segment_distance2(f.capA, f.u, f.a, c2a, c2b) < P.k.contact2) {
                                unsigned bit = 16;
#pragma unroll
                                for (int i = 0; i < GROUP; i++) bit = jq == kIdx[i] ? 1u << i : bit;
                                hitAll |= bit << (5 * k);
                            }
                        }
                        __builtin_amdgcn_wave_barrier();
                        queued = 0;
                        if (clk) tPassB += wall_clock64() - tb0;
                    }
                }
                if (hitAll) atomicOr(&hitWords[buf], (int)hitAll);
            }
            const unsigned long long tg1 = clk ? wall_clock64() : 0;
            __syncthreads();  // the group's only barrier: hit words and the next plan are complete
            const unsigned long long tg2 = clk ? wall_clock64() : 0;
            // decide the group's moves in order (uniform: every lane of every wavefront reaches the same decisions)
            // move k is rejected if it leaves the cell, overlaps a particle outside the group, overlaps the old
            // pose of a group particle i not moved before it (bit i of the stencil hits: i > k, or i < k rejected),
            // or the trial pose of an accepted move i < k (bit i of the planner's trial-pair hits); bit k of the
            // stencil hits is never set (a move is not tested against its own particle)
            const unsigned hw = (unsigned)__builtin_amdgcn_readfirstlane(hitWords[buf]);
            const unsigned word0 = word, newHit0 = newHit;
            unsigned acc = 0, insideMask = 0;
            for (int k = 0; k < gs; k++) {
                const unsigned b = (hw >> (5 * k)) & 31;
                const unsigned inside = (word0 >> (4 * k + 3)) & 1;
                const unsigned trialHits = (newHit0 >> (4 * k)) & 15;
                const bool hit = !inside || (b & 16) || ((b & 15) & ~acc) || (trialHits & acc);
                insideMask |= inside << k;
                if (!hit) acc |= 1u << k;
            }
            accepted += (unsigned long long)__popc(acc);
            leftCell += (unsigned long long)(gs - __popc(insideMask));
            tests += (unsigned long long)total * (unsigned long long)__popc(insideMask);
            if (G32 && !planner && lane == 0 && wave == 1)  // pairs the scan classified (the moved particles' own
                g32Pairs += (unsigned long long)(total - 1) * (unsigned long long)__popc(insideMask);  // excluded)
            const bool replan = (acc & nextUses) != 0;
            // accepted poses into this wavefront's view of the staged copy, one lane per component; LDS accesses of
            // a wavefront complete in order, so its next scan reads them
            for (int k = 0; k < gs; k++) {
                if (!((acc >> k) & 1)) continue;
                const int ki = kIdx[k];
                const double *q = plan + k * PLAN_D;
                if (lane < (KIND == 1 ? 6 : 3)) {
                    double *dst = lane == 0 ? S.x : lane == 1 ? S.y : lane == 2 ? S.z
                                : lane == 3 ? S.ax : lane == 4 ? S.ay : S.az;
                    dst[ki] = q[lane];
                } else if (lane == 6) {
                    S.code[ki] = wrap_code((word0 >> (4 * k)) & 7);
                } else if (lane == 7 && planner) {
                    // marks the move accepted: a decided move's particle index is not read again in this chunk
                    pk[m0 + k] |= PK_ACCEPTED;
                } else if (G32 && lane >= 17 && lane < 20) {
                    Sf[(lane - 17) * P.cap + ki] = reinterpret_cast<const float *>(q + 22)[lane - 17];
                } else if (planner && lane >= 8 && lane < 17) {
                    oo[(lane - 8) * OWN_CAP + ki] = q[13 + (lane - 8)];
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (replan) {
                // the next group was planned from a pose this group just replaced
                __syncthreads();
                if (planner)
                    plan_group<KIND, TRANSLATED, G32 != 0>(P, S, oo, pt, pR, pk, sT, m1, gsNext, cell,
                                                           plans + nbuf * GROUP * PLAN_D, plansI + nbuf * GPLAN_I, o);
                __syncthreads();
            }
            if (clk) {
                tBusy += tg1 - tg0;
                tDecide += wall_clock64() - tg2;
                tTop += tg0 - tgTop;
                nGroups++;
                nReplans += replan ? 1 : 0;
            }
        }
        __syncthreads();
        // per-proposal acceptance flags (several samplers' counters, decision records); each lane reads the entry
        // it alone rewrites with the next chunk's proposals
        if (P.pr.acc && tid < cn && (pk[tid] & PK_ACCEPTED)) P.pr.acc[pbase + chunk + tid] = 1;
    }
    if (clk) {
        clk[3] = wall_clock64();
        if (lane == 0 && wave <= 1) atomicAdd(&P.phases[planner ? RP_PHASE_PLAN : RP_PHASE_SCAN], tBusy);
        if (lane == 0 && wave == 1) atomicAdd(&P.phases[RP_PHASE_PASSB], tPassB);
        if (tid == 0) {
            atomicAdd(&P.phases[RP_PHASE_DECIDE], tDecide);
            atomicAdd(&P.phases[RP_PHASE_GROUPS], nGroups);
            atomicAdd(&P.phases[RP_PHASE_REPLANS], nReplans);
            atomicAdd(&P.phases[RP_PHASE_TOP], tTop);
        }
    }
    for (int a = tid; a < nOwn; a += WAVE * WAVES) {
        const int gi = own0 + a;
        P.p.x[gi] = S.x[a];
        P.p.y[gi] = S.y[a];
        P.p.z[gi] = S.z[a];
        for (int c = 0; c < 9; c++) P.p.o[c][gi] = oo[c * OWN_CAP + a];
        int code = S.code[a];
        P.p.wbits[gi] = (uint8_t)(((code & 3) - 1) | ((((code >> 2) & 3) - 1) << 1) | ((((code >> 4) & 3) - 1) << 2));
    }
    flush_counters(P, cell, accepted, moves, leftCell, tests, true);
    if (G32 && !planner) {
        // per-wavefront sums of the lanes' statistics, one atomic each
        for (int off = WAVE / 2; off > 0; off >>= 1) {
            g32Guard += __shfl_xor(g32Guard, off);
            g32Mismatch += __shfl_xor(g32Mismatch, off);
            g32Err = fmaxf(g32Err, __shfl_xor(g32Err, off));
        }
        if (lane == 0) {
            if (wave == 1) P.slots[(size_t)cell * SLOT_WORDS + SLOT_G32_PAIRS] = g32Pairs;
            if (g32Guard) atomicAdd(&P.counters[CNT_G32_GUARD], (unsigned long long)g32Guard);
            if (G32 == 2) {
                if (g32Mismatch) atomicAdd(&P.counters[CNT_G32_MISMATCH], (unsigned long long)g32Mismatch);
                atomicMax(&P.counters[CNT_G32_MAXERR], (unsigned long long)__float_as_uint(g32Err));
            }
        }
    }
}

// Soft sweep of one active cell (Packing::tryMove with a soft interaction, Packing.cpp:136-164, and the Metropolis
// test of Simulation.cpp:483).  In overlap-counting mode (Packing::calculateMoveOverlapEnergy, Packing.cpp:304-323)
// the summed pair value is the overlap indicator, the move is accepted if it stays in the cell and its overlap
// count does not grow, and the accepted changes are summed into counters[CNT_OVERLAP_DELTA].  Every trial move sums the moved particle's energy with all stencil candidates at
// its current pose (initialEnergy) and at the trial pose (finalEnergy) in one pass; the lanes' sums are reduced per
// wavefront (xor butterfly: every lane ends with the same bits), the wavefronts' sums in index order, and the move is
// accepted if it stays in the cell and u <= exp(-(finalEnergy - initialEnergy) / T).  STAGED: stencil and own
// orientations in LDS; otherwise read from HBM and accepted poses written back there by lane 0.
static_assert(SWEEP_AUX >= 2 * PLAN_D && SWEEP_AUX >= MOVE_CHUNK + 2 * 2 * 4, "sweep scratch area too small");

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <int KIND, int WAVES, bool TRANSLATED, bool STAGED>
__device__ __forceinline__ void sweep_cell_soft(const SweepParams &P, const Stencil &S, double *oo, double *pt, double *pR, int *pk,
                                double *aux, const double *sT, int cell, int own0, int nOwn, int total) {
    constexpr int NT = WAVE * WAVES;
    const Grid &g = P.g;
    const int tid = threadIdx.x;
    const int wave = tid / WAVE;
    double *pu = aux;                 // MOVE_CHUNK uniforms
    double *sums = aux + MOVE_CHUNK;  // [move parity][final, initial][wave]
    unsigned long long accepted = 0, leftCell = 0, tests = 0;
    long long overlapDelta = 0;
    int move = 0;
    const int moves = P.moves > 0 ? P.moves : nOwn * P.mult;
    const int pbase = P.moves > 0 ? cell * P.moves : own0 * P.mult;
    // polyspheres (counting mode): the bead table, the moved particle's bead centres before and after the move
    __shared__ double sBeads[KIND == 2 ? 4 * RP_MAX_CENTRES : 1], sOld[KIND == 2 ? 3 * RP_MAX_CENTRES : 1],
        sNew[KIND == 2 ? 3 * RP_MAX_CENTRES : 1];
    if (KIND == 2) {
        for (int e = tid; e < 4 * P.k.nc; e += NT) sBeads[e] = P.k.beads[e];
        __syncthreads();
    }
    for (int chunk = 0; chunk < moves; chunk += MOVE_CHUNK) {
        const int cn = min(MOVE_CHUNK, moves - chunk);
        load_proposals(P, pbase, chunk, cn, pt, pR, pk);
        if (tid < cn && P.pr.u) pu[tid] = P.pr.u[pbase + chunk + tid];
        __syncthreads();
        for (int mi = 0; mi < cn; mi++, move++) {
            const int kIdx = pk[mi];
            const int gi = own0 + kIdx;
            V3 opos, oax;
            double O[9];
            int ow[3];
            if (STAGED) {
                opos = v3(S.x[kIdx], S.y[kIdx], S.z[kIdx]);
                for (int c = 0; c < 9; c++) O[c] = oo[c * OWN_CAP + kIdx];
                const int code = S.code[kIdx];
                ow[0] = (code & 3) - 1;
                ow[1] = ((code >> 2) & 3) - 1;
                ow[2] = ((code >> 4) & 3) - 1;
            } else {
                opos = load3(P.p.x, P.p.y, P.p.z, gi);
                for (int c = 0; c < 9; c++) O[c] = P.p.o[c][gi];
                const unsigned w = P.p.wbits[gi];
                ow[0] = w & 1;
                ow[1] = (w >> 1) & 1;
                ow[2] = (w >> 2) & 1;
            }
            oax = v3(O[2], O[5], O[8]);
            V3 t = v3(pt[mi], pt[MOVE_CHUNK + mi], pt[2 * MOVE_CHUNK + mi]);
            V3 npos = translate_wrapped(g.dims, g.inv, opos, t);
            int nc[3];
            unsigned nw;
            bin_position(g, npos, nc, nw);
            const bool inside = cell_index(g, nc[0], nc[1], nc[2]) == cell;
            double NO[9], R[9];
            for (int c = 0; c < 9; c++) R[c] = pR[c * MOVE_CHUNK + mi];
            mulm(R, O, NO);  // Shape::rotate (Shape.cpp:14-16)
            const V3 nax = v3(NO[2], NO[5], NO[8]);
            const int nwx = nw & 1, nwy = (nw >> 1) & 1, nwz = (nw >> 2) & 1;
            double eFinal = 0, eInitial = 0;
            if (KIND == 2) {
                // bead centres pos + O c before and after the move (PolysphereTraits.cpp:93-95)
                if (tid < P.k.nc) {
                    const V3 c = v3(sBeads[4 * tid], sBeads[4 * tid + 1], sBeads[4 * tid + 2]);
                    const V3 cn2 = add(npos, mulv(NO, c)), co = add(opos, mulv(O, c));
                    sNew[3 * tid] = cn2.x; sNew[3 * tid + 1] = cn2.y; sNew[3 * tid + 2] = cn2.z;
                    sOld[3 * tid] = co.x; sOld[3 * tid + 1] = co.y; sOld[3 * tid + 2] = co.z;
                }
                __syncthreads();
            }
            if (inside) {
                SCFirst fN = make_first<KIND>(P.k, npos, nax);
                SCFirst fO = make_first<KIND>(P.k, opos, oax);
                for (int j = tid; j < total; j += NT) {
                    if (j == kIdx) continue;
                    // unstaged stencils may hold the own cell again under another image (one cell across an
                    // axis): the moved particle is skipped there too (Packing.cpp:383-384)
                    if (!STAGED) {
                        const int q = stencil_cell(S.start, S.ncells, j);
                        if (S.base[q] + (j - S.start[q]) == gi) continue;
                    }
                    V3 p2, a2;
                    int code;
                    load_candidate<KIND, STAGED>(P.p, S, j, p2, a2, code);
                    V3 TN = v3(0, 0, 0), TO = v3(0, 0, 0);
                    if (!TRANSLATED) {
                        TN = image_translation(sT, code, nwx, nwy, nwz);
                        TO = image_translation(sT, code, ow[0], ow[1], ow[2]);
                    }
                    if (KIND == 2) {
                        const int q = stencil_cell(S.start, S.ncells, j);
                        const int gj = S.base[q] + (j - S.start[q]);
                        eFinal += (double)poly_count(P.k, sBeads, sNew, npos, p2, P.p, gj, TN);
                        eInitial += (double)poly_count(P.k, sBeads, sOld, opos, p2, P.p, gj, TO);
                    } else if (P.counting) {
                        eFinal += pair_overlap<KIND>(P.k, fN, p2, a2, TN) ? 1.0 : 0.0;
                        eInitial += pair_overlap<KIND>(P.k, fO, p2, a2, TO) ? 1.0 : 0.0;
                    } else {
                        eFinal += soft_energy(P.k, fN, p2, a2, TN);
                        eInitial += soft_energy(P.k, fO, p2, a2, TO);
                        // CompoundInteraction: an overlap of the hard parts makes the move's energy infinite
                        // (Packing::calculateMoveOverlapEnergy, Packing.cpp:304-323)
                        if (P.k.hardCore && pair_overlap<KIND>(P.k, fN, p2, a2, TN)) eFinal = INFINITY;
                    }
                }
                tests += (unsigned long long)total;
            } else {
                leftCell++;
            }
            eFinal = wave_sum(eFinal);
            eInitial = wave_sum(eInitial);
            double *slot = sums + (move & 1) * 2 * WAVES;
            if ((tid & (WAVE - 1)) == 0) {
                slot[wave] = eFinal;
                slot[WAVES + wave] = eInitial;
            }
            __syncthreads();  // one barrier per move: the two slots alternate
            double fin = 0, ini = 0;
            for (int w = 0; w < WAVES; w++) {
                fin += slot[w];
                ini += slot[WAVES + w];
            }
            if (!inside) continue;
            if (P.counting) {
                if (fin - ini > 0) continue;  // counts are exact in doubles
                overlapDelta += (long long)(fin - ini);
            } else if (!(fin < INFINITY) || !(pu[mi] <= exp(-(fin - ini) / P.temperature))) {
                // an overlap rejects the move outright (u = 0 would pass u <= exp(-inf))
                continue;
            }

            accepted++;
            if (WAVES < 8 && P.pr.acc && tid == 0) P.pr.acc[pbase + chunk + mi] = 1;
            if (STAGED) {
                S.x[kIdx] = npos.x;
                S.y[kIdx] = npos.y;
                S.z[kIdx] = npos.z;
                if (KIND == 1) {
                    S.ax[kIdx] = nax.x;
                    S.ay[kIdx] = nax.y;
                    S.az[kIdx] = nax.z;
                }
                S.code[kIdx] = (nwx + 1) | ((nwy + 1) << 2) | ((nwz + 1) << 4);
                for (int c = 0; c < 9; c++) oo[c * OWN_CAP + kIdx] = NO[c];
            } else {
                if (tid == 0) {
                    P.p.x[gi] = npos.x;
                    P.p.y[gi] = npos.y;
                    P.p.z[gi] = npos.z;
                    for (int c = 0; c < 9; c++) P.p.o[c][gi] = NO[c];
                    P.p.wbits[gi] = (uint8_t)nw;
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
    if (STAGED) {
        for (int a = tid; a < nOwn; a += NT) {
            const int gi = own0 + a;
            P.p.x[gi] = S.x[a];
            P.p.y[gi] = S.y[a];
            P.p.z[gi] = S.z[a];
            for (int c = 0; c < 9; c++) P.p.o[c][gi] = oo[c * OWN_CAP + a];
            int code = S.code[a];
            P.p.wbits[gi] =
                (uint8_t)(((code & 3) - 1) | ((((code >> 2) & 3) - 1) << 1) | ((((code >> 4) & 3) - 1) << 2));
        }
    }
    flush_counters(P, cell, accepted, moves, leftCell, tests, STAGED);
    if (P.counting && tid == 0 && overlapDelta)
        atomicAdd(&P.counters[CNT_OVERLAP_DELTA], (unsigned long long)overlapDelta);
}

// Stages the stencil: every lane resolves up to SLOT_GROUP entries, issues all their HBM loads, then writes LDS.
// With G32 each entry's position relative to o is also staged in FP32 (fx[j], fx[cap + j], fx[2 cap + j]): the
// FP64 difference rounded once to FP32.
template <int KIND, int NT, bool G32 = false>
__device__ __forceinline__ void stage_stencil_batched(const SweepParams &P, const Stencil &S, const double *sT,
                                                      int total, bool translated, float *fx = nullptr,
                                                      V3 o = V3{}) {
    const int tid = threadIdx.x;
    for (int base = 0; base < total; base += NT * SLOT_GROUP) {
        int gi[SLOT_GROUP], code[SLOT_GROUP];
#pragma unroll
        for (int s = 0; s < SLOT_GROUP; s++) {
            const int j = base + s * NT + tid;
            gi[s] = -1;
            if (j < total) {
                int q = stencil_cell(S.start, S.ncells, j);
                gi[s] = S.base[q] + (j - S.start[q]);
                code[s] = S.ecode[q];
            }
        }
        V3 p[SLOT_GROUP], a[SLOT_GROUP];
        int w[SLOT_GROUP];
#pragma unroll
        for (int s = 0; s < SLOT_GROUP; s++) {
            if (gi[s] >= 0) {
                p[s] = load3(P.p.x, P.p.y, P.p.z, gi[s]);
                if (KIND == 1) a[s] = v3(P.p.o[2][gi[s]], P.p.o[5][gi[s]], P.p.o[8][gi[s]]);
                w[s] = P.p.wbits[gi[s]];
            }
        }
#pragma unroll
        for (int s = 0; s < SLOT_GROUP; s++) {
            const int j = base + s * NT + tid;
            if (gi[s] >= 0) {
                const int c = image_code(code[s], w[s]);
                V3 q = translated ? add(p[s], image_translation(sT, c, 0, 0, 0)) : p[s];
                S.x[j] = q.x;
                S.y[j] = q.y;
                S.z[j] = q.z;
                if (G32) {
                    fx[j] = (float)(q.x - o.x);
                    fx[P.cap + j] = (float)(q.y - o.y);
                    fx[2 * P.cap + j] = (float)(q.z - o.z);
                }
                if (KIND == 1) {
                    S.ax[j] = a[s].x;
                    S.ay[j] = a[s].y;
                    S.az[j] = a[s].z;
                }
                S.code[j] = c;
            }
        }
    }
}

// Active cell b of colour P.colour (one workgroup).
// LEAN != 0: the kernel carries only the grouped sweep of staged cells (and the HBM path for cells that cannot be
// staged), with fewer live registers than the general kernel: 1 FP64 scan, 2 guarded FP32 scan, 3 guarded FP32 scan
// with the FP64 verification of every pair.
template <int KIND, int WAVES, bool SOFT, int LEAN>
__device__ __forceinline__ void sweep_block(const SweepParams &P, const int b) {
    constexpr int G32 = LEAN >= 2 ? LEAN - 1 : 0;
    constexpr int NT = WAVE * WAVES;
    extern __shared__ double smem[];
    __shared__ int s_start[MAX_STENCIL + 1];
    __shared__ int s_base[MAX_STENCIL];
    __shared__ int s_ecode[MAX_STENCIL];
    __shared__ int s_vote[2];

    const Grid &g = P.g;
    const int tid = threadIdx.x;
    const unsigned long long t0 = P.trace ? wall_clock64() : 0;
    const int bx = b % P.na[0], by = (b / P.na[0]) % P.na[1], bz = b / (P.na[0] * P.na[1]);
    const int per = g.period;
    const int cx = per * bx + P.colour % per, cy = per * by + (P.colour / per) % per,
              cz = per * bz + P.colour / (per * per);
    if (cx >= g.n[0] || cy >= g.n[1] || cz >= g.n[2]) return;  // odd parity on a one-cell axis
    const int cell = cell_index(g, cx, cy, cz);
    const int own0 = g.cellStart[cell];
    const int nOwn = g.cellStart[cell + 1] - own0;
    if (nOwn == 0) return;
    unsigned long long clk[5];
    if (P.phases) clk[0] = wall_clock64();

    const int ncells = stencil_cells(g.reach, false);
    if (tid < 2) s_vote[tid] = 0;
    // polyspheres reject most moves on a near neighbour: their stencil is searched nearest cells first
    build_stencil(g, cx, cy, cz, ncells, false, s_start, s_base, s_ecode, KIND == 2);
    const int total = s_start[ncells];
    if (P.phases) clk[1] = wall_clock64();

    Stencil S;
    S.start = s_start;
    S.base = s_base;
    S.ecode = s_ecode;
    S.ncells = ncells;
    double *rest;
    carve_stencil(S, smem, P.cap, KIND, rest);
    double *oo = rest;                  // 9 x OWN_CAP own orientations
    double *pt = oo + 9 * OWN_CAP;      // 3 x MOVE_CHUNK proposal translations
    double *pR = pt + 3 * MOVE_CHUNK;   // 9 x MOVE_CHUNK proposal rotations
    double *sT = pR + 9 * MOVE_CHUNK;   // 81 translation table
    double *plans = sT + 81;            // SWEEP_AUX: 2 x PLAN_D planner output, or soft uniforms and sums
    int *pk = reinterpret_cast<int *>(plans + SWEEP_AUX);
    int *plansI = pk + MOVE_CHUNK;      // SWEEP_AUX_I planner output ints
    int *queue = plansI + SWEEP_AUX_I;  // max(cap + 64*WAVES, (WAVES-1)*GROUPED_QUEUE) queued candidates
    // G32: the FP32 stencil positions (3 x cap) and translation table (81) after the queue
    float *Sf = reinterpret_cast<float *>(queue + max(P.cap + WAVE * WAVES, (WAVES - 1) * GROUPED_QUEUE));
    float *sTf = Sf + 3 * P.cap;
    for (int i = tid; i < 81; i += NT) {
        sT[i] = g.T[i];
        if (G32) sTf[i] = (float)g.T[i];
    }
    __syncthreads();

    // with one cell across an axis the own cell recurs in its stencil under other images; staged copies would go
    // stale as particles move, so such grids read the stencil from HBM
    const bool ownRecurs = g.n[0] == 1 || g.n[1] == 1 || g.n[2] == 1;
    if (LEAN && KIND != 2 && total <= P.cap && nOwn <= OWN_CAP && !ownRecurs) {
        // grouped moves only; the guarded FP32 scan takes positions relative to the cell's centre
        const bool translated = cx < g.n[0] - 1 && cy < g.n[1] - 1 && cz < g.n[2] - 1;
        const V3 o = G32 ? mulv(g.dims, v3(g.shift[0] + ((double)cx + 0.5) * g.relCell[0],
                                           g.shift[1] + ((double)cy + 0.5) * g.relCell[1],
                                           g.shift[2] + ((double)cz + 0.5) * g.relCell[2]))
                         : v3(0, 0, 0);
        stage_stencil_batched<KIND, NT, G32 != 0>(P, S, sT, total, translated, Sf, o);
        for (int e = tid; e < 9 * nOwn; e += NT) {
            int c = e / nOwn, a = e - c * nOwn;
            oo[c * OWN_CAP + a] = P.p.o[c][own0 + a];
        }
        __syncthreads();
        int *hitWords = plansI + GBUF * GPLAN_I;
        if (translated)
            sweep_cell_grouped<KIND, WAVES, true, G32>(P, S, oo, pt, pR, pk, queue, sT, plans, plansI, hitWords, own0,
                                                       nOwn, total, cell, nullptr, Sf, sTf, o);
        else
            sweep_cell_grouped<KIND, WAVES, false, G32>(P, S, oo, pt, pR, pk, queue, sT, plans, plansI, hitWords, own0,
                                                        nOwn, total, cell, nullptr, Sf, sTf, o);
    } else if (LEAN) {
        sweep_cell_unstaged<KIND, WAVES>(P, S, pt, pR, pk, sT, cell, own0, nOwn, total);
    } else if (KIND != 2 && total <= P.cap && nOwn <= OWN_CAP && !ownRecurs) {
        // a cell off the last row on every axis holds no wrapped particle: stage pos + T directly
        const bool translated = cx < g.n[0] - 1 && cy < g.n[1] - 1 && cz < g.n[2] - 1;
        stage_stencil_batched<KIND, NT>(P, S, sT, total, translated);
        for (int e = tid; e < 9 * nOwn; e += NT) {
            int c = e / nOwn, a = e - c * nOwn;
            oo[c * OWN_CAP + a] = P.p.o[c][own0 + a];
        }
        __syncthreads();
        if (P.phases) clk[2] = wall_clock64();
        unsigned long long *c3 = P.phases ? clk : nullptr;
        if (SOFT) {
            if (translated)
                sweep_cell_soft<KIND, WAVES, true, true>(P, S, oo, pt, pR, pk, plans, sT, cell, own0, nOwn, total);
            else
                sweep_cell_soft<KIND, WAVES, false, true>(P, S, oo, pt, pR, pk, plans, sT, cell, own0, nOwn, total);
            if (c3) c3[3] = wall_clock64();
        } else if (P.fixMode) {
            if (translated)
                sweep_cell_staged<KIND, WAVES, true>(P, S, oo, pt, pR, pk, queue, sT, s_vote, cell, own0, nOwn, total, c3);
            else
                sweep_cell_staged<KIND, WAVES, false>(P, S, oo, pt, pR, pk, queue, sT, s_vote, cell, own0, nOwn, total, c3);
        } else if (WAVES >= 2 && P.pipelined == 2) {
            int *hitWords = plansI + GBUF * GPLAN_I;
            if (translated)
                sweep_cell_grouped<KIND, WAVES, true>(P, S, oo, pt, pR, pk, queue, sT, plans, plansI, hitWords, own0,
                                                      nOwn, total, cell, c3);
            else
                sweep_cell_grouped<KIND, WAVES, false>(P, S, oo, pt, pR, pk, queue, sT, plans, plansI, hitWords, own0,
                                                       nOwn, total, cell, c3);
        } else if (WAVES >= 2 && P.pipelined) {
            if (translated)
                sweep_cell_pipelined<KIND, WAVES, true>(P, S, oo, pt, pR, pk, queue, sT, s_vote, plans, plansI, own0,
                                                        nOwn, total, cell, c3);
            else
                sweep_cell_pipelined<KIND, WAVES, false>(P, S, oo, pt, pR, pk, queue, sT, s_vote, plans, plansI, own0,
                                                         nOwn, total, cell, c3);
        } else if (translated) {
            sweep_cell_staged<KIND, WAVES, true>(P, S, oo, pt, pR, pk, queue, sT, s_vote, cell, own0, nOwn, total, c3);
        } else {
            sweep_cell_staged<KIND, WAVES, false>(P, S, oo, pt, pR, pk, queue, sT, s_vote, cell, own0, nOwn, total, c3);
        }
    } else {
        if (P.phases) clk[2] = wall_clock64();
        if (SOFT)
            sweep_cell_soft<KIND, WAVES, false, false>(P, S, oo, pt, pR, pk, plans, sT, cell, own0, nOwn, total);
        else
            sweep_cell_unstaged<KIND, WAVES>(P, S, pt, pR, pk, sT, cell, own0, nOwn, total);
        if (P.phases) clk[3] = wall_clock64();
    }
    if (P.phases && tid == 0) {
        clk[4] = wall_clock64();
        atomicAdd(&P.phases[RP_PHASE_TABLE], clk[1] - clk[0]);
        atomicAdd(&P.phases[RP_PHASE_STAGE], clk[2] - clk[1]);
        atomicAdd(&P.phases[RP_PHASE_MOVES], clk[3] - clk[2]);
        atomicAdd(&P.phases[RP_PHASE_WRITE], clk[4] - clk[3]);
        atomicAdd(&P.phases[RP_PHASE_BLOCKS], 1ull);
        atomicMax(&P.phases[RP_PHASE_LAST], clk[4] - clk[0]);
    }
    if (P.trace && tid == 0) {
        unsigned long long *t = P.trace + 4 * b;
        t[0] = t0;
        t[1] = wall_clock64();
        t[2] = (unsigned long long)nOwn | ((unsigned long long)total << 16) | ((unsigned long long)cell << 32);
        t[3] = (unsigned long long)__smid();
    }
}

// Two workgroups of eight wavefronts share a compute unit only within 128 registers per lane (4 wavefronts per SIMD)
template <int KIND, int WAVES, bool SOFT, int LEAN = 0>
__global__ __launch_bounds__(WAVE *WAVES, WAVES >= 8 ? 4 : 1) void k_sweep(SweepParams P) {
    sweep_block<KIND, WAVES, SOFT, LEAN>(P, (int)blockIdx.x);
}

// One colour pass of several packings in one launch: packing r owns workgroups [first[r], first[r + 1]) and its
// parameters (grid, particles, proposals, counters, colour) are table[r]; every workgroup runs exactly the sweep of
// that packing's active cell b - first[r].
template <int KIND, int WAVES, int LEAN>
__global__ __launch_bounds__(WAVE *WAVES, WAVES >= 8 ? 4 : 1) void k_sweep_batch(
    const SweepParams *__restrict__ table, const int *__restrict__ first, int packings) {
    const int b = (int)blockIdx.x;
    int lo = 0, hi = packings;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (first[mid] <= b) lo = mid; else hi = mid;
    }
    // the table is read-only for the launch: read through the constant address space, its fields load as uniform
    // scalars, as the single-packing kernel's argument does
    typedef const __attribute__((address_space(4))) SweepParams ConstParams;
    ConstParams *entry = (ConstParams *)table + lo;
    sweep_block<KIND, WAVES, false, LEAN>(*(const SweepParams *)entry, b - first[lo]);
}

// ---------------------------------------------------------------------------------------------------------------
// Candidate pairs of the learned ordering (polyspheres).  Move q: a uniform pick among the n sorted records and a
// rototranslation drawn as k_proposals draws one (translation half-width ts, rotation half-angle rs), from its own
// Philox stream (seed, stream, q); a trial pose that leaves its checkerboard cell is never tested and has no pairs.
// The close candidates of a trial pose -- those within the total interaction range under the stencil's image
// translation (poly_hit's first test), the particle itself skipped -- in natural stencil order (the own cell, then
// the others in lexicographic order of their offsets; records in storage order within a cell), each with
//   12 features: the candidate's position in the trial body frame divided by the total range (NO^T (p2 + T - npos)),
//                and the relative rotation NO^T O_j (row-major),
//   the exact decision poly_hit (1 = the bead pairs overlap),
//   meta: move q, its rank in natural order, and its key in the nearest-first order (shell position of its cell
//         << 20 | rank within the cell).
// mode 0 counts the close candidates of every move (counts[q]); mode 1 writes them from offsets[q] on.
constexpr int MLP_PAIR_THREADS = 64;

__global__ __launch_bounds__(MLP_PAIR_THREADS) void k_mlp_pairs(SweepParams P, int M, uint64_t seed, uint32_t stream,
                                                                double ts, double rs, int mode, int *counts,
                                                                const long long *offsets, float *x, float *y,
                                                                int *meta) {
    __shared__ double sT[81];
    __shared__ double sBeads[4 * RP_MAX_CENTRES];
    __shared__ double sTrialAll[MLP_PAIR_THREADS][3 * RP_MAX_CENTRES];
    __shared__ unsigned char sShellPos[MAX_STENCIL];
    const Grid &g = P.g;
    const int tid = threadIdx.x;
    for (int i = tid; i < 81; i += blockDim.x) sT[i] = g.T[i];
    for (int i = tid; i < 4 * P.k.nc; i += blockDim.x) sBeads[i] = P.k.beads[i];
    for (int i = tid; i < MAX_STENCIL; i += blockDim.x) sShellPos[kShellOrder[i]] = (unsigned char)i;
    __syncthreads();
    const int q = blockIdx.x * blockDim.x + tid;
    if (q >= M) return;
    double *sTrial = sTrialAll[tid];
    const int n = g.cellStart[g.ncell];
    Philox rng = philox_init(seed ^ 0x6d6c7061697273ull, stream, 0xfffffffeu, (uint32_t)q);
    const int gi = (int)philox_below(rng, (uint32_t)n);
    V3 t;
    t.x = philox_uniform(rng, -ts, ts);
    t.y = philox_uniform(rng, -ts, ts);
    t.z = philox_uniform(rng, -ts, ts);
    V3 axis;
    do {
        axis.x = philox_uniform(rng, -1, 1);
        axis.y = philox_uniform(rng, -1, 1);
        axis.z = philox_uniform(rng, -1, 1);
    } while (dot(axis, axis) > 1 || dot(axis, axis) == 0);
    double R[9], O[9], NO[9];
    rodrigues(axis, philox_uniform(rng, -rs, rs), R);
    for (int c = 0; c < 9; c++) O[c] = P.p.o[c][gi];
    mulm(R, O, NO);
    const V3 npos = translate_wrapped(g.dims, g.inv, load3(P.p.x, P.p.y, P.p.z, gi), t);
    int nc[3];
    unsigned nw;
    bin_position(g, npos, nc, nw);
    const int cell = P.p.cell[gi];
    int found = 0;
    if (cell_index(g, nc[0], nc[1], nc[2]) == cell) {
        for (int a = 0; a < P.k.nc; a++) {
            const V3 c1 = add(npos, mulv(NO, v3(sBeads[4 * a], sBeads[4 * a + 1], sBeads[4 * a + 2])));
            sTrial[3 * a] = c1.x;
            sTrial[3 * a + 1] = c1.y;
            sTrial[3 * a + 2] = c1.z;
        }
        const int cx = cell % g.n[0], cy = (cell / g.n[0]) % g.n[1], cz = cell / (g.n[0] * g.n[1]);
        const int k = g.reach, side = 2 * k + 1, centre = (side * side * side - 1) / 2;
        const int ncells = side * side * side;
        const int wx = nw & 1, wy = (nw >> 1) & 1, wz = (nw >> 2) & 1;
        const long long out0 = mode == 1 ? offsets[q] : 0;
        for (int slot = 0; slot < ncells; slot++) {
            const int tt = slot == 0 ? centre : (slot <= centre ? slot - 1 : slot);
            int ex, ey, ez;
            const int mx = wrap_axis(cx, tt / (side * side) - k, g.n[0], ex);
            const int my = wrap_axis(cy, (tt / side) % side - k, g.n[1], ey);
            const int mz = wrap_axis(cz, tt % side - k, g.n[2], ez);
            const int mc = cell_index(g, mx, my, mz);
            const int ecode = (ex + 1) | ((ey + 1) << 2) | ((ez + 1) << 4);
            const int b0 = g.cellStart[mc], b1 = g.cellStart[mc + 1];
            for (int gj = b0; gj < b1; gj++) {
                if (gj == gi) continue;
                const V3 p2 = load3(P.p.x, P.p.y, P.p.z, gj);
                const V3 T = image_translation(sT, image_code(ecode, P.p.wbits[gj]), wx, wy, wz);
                const V3 d = sub(add(p2, T), npos);
                if (!(dot(d, d) < P.k.range2)) continue;
                if (mode == 1) {
                    const long long o = out0 + found;
                    double Oj[9];
                    for (int c = 0; c < 9; c++) Oj[c] = P.p.o[c][gj];
                    float *f = x + o * MLP_FEATURES;
                    // NO^T d and NO^T O_j
                    for (int r = 0; r < 3; r++)
                        f[r] = (float)((NO[r] * d.x + NO[3 + r] * d.y + NO[6 + r] * d.z) * P.mlpInvRange);
                    for (int r = 0; r < 3; r++)
                        for (int c = 0; c < 3; c++)
                            f[3 + 3 * r + c] = (float)(NO[r] * Oj[c] + NO[3 + r] * Oj[3 + c] + NO[6 + r] * Oj[6 + c]);
                    y[o] = poly_hit(P.k, sBeads, sTrial, npos, p2, P.p, gj, T) ? 1.f : 0.f;
                    meta[3 * o] = q;
                    meta[3 * o + 1] = found;
                    meta[3 * o + 2] = ((int)sShellPos[tt] << 20) | (gj - b0);
                }
                found++;
            }
        }
    }
    if (mode == 0) counts[q] = found;
}

hipError_t launch_mlp_pairs(hipStream_t s, const SweepParams &sp, int M, uint64_t seed, uint32_t stream, double ts,
                            double rs, int mode, int *counts, const long long *offsets, float *x, float *y,
                            int *meta) {
    if (M <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_pairs, dim3((unsigned)((M + MLP_PAIR_THREADS - 1) / MLP_PAIR_THREADS)),
                       dim3(MLP_PAIR_THREADS), 0, s, sp, M, seed, stream, ts, rs, mode, counts, offsets, x, y, meta);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------------------
// Nematic Q-tensor sums (NematicOrder::calculate, NematicOrder.cpp:9-17): sum over particles of a_i a_j for the
// primary axis a = orientation column 2 (SpherocylinderTraits.cpp:52-54).  Each product is rounded once, as in the
// reference; the sum runs as a fixed tree (block partials in a fixed order, then one workgroup), so a given state
// always gives the same bits.

constexpr int NEMATIC_THREADS = 256;
constexpr int NEMATIC_BLOCKS = 256;

__device__ __forceinline__ void block_sum6(double v[6], double *lds) {
    const int tid = threadIdx.x;
    for (int c = 0; c < 6; c++) lds[c * NEMATIC_THREADS + tid] = v[c];
    __syncthreads();
    for (int w = NEMATIC_THREADS / 2; w > 0; w >>= 1) {
        if (tid < w)
            for (int c = 0; c < 6; c++) lds[c * NEMATIC_THREADS + tid] += lds[c * NEMATIC_THREADS + tid + w];
        __syncthreads();
    }
}

__global__ __launch_bounds__(NEMATIC_THREADS) void k_nematic_partial(Particles p, int n, double *partial) {
    __shared__ double lds[6 * NEMATIC_THREADS];
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * NEMATIC_THREADS + threadIdx.x; i < n; i += NEMATIC_BLOCKS * NEMATIC_THREADS) {
        const double a0 = p.o[2][i], a1 = p.o[5][i], a2 = p.o[8][i];
        v[0] += a0 * a0;
        v[1] += a0 * a1;
        v[2] += a0 * a2;
        v[3] += a1 * a1;
        v[4] += a1 * a2;
        v[5] += a2 * a2;
    }
    block_sum6(v, lds);
    if (threadIdx.x < 6) partial[threadIdx.x * NEMATIC_BLOCKS + blockIdx.x] = lds[threadIdx.x * NEMATIC_THREADS];
}

__global__ __launch_bounds__(NEMATIC_THREADS) void k_nematic_final(const double *partial, double *out) {
    __shared__ double lds[6 * NEMATIC_THREADS];
    double v[6];
    for (int c = 0; c < 6; c++) v[c] = partial[c * NEMATIC_BLOCKS + threadIdx.x];
    block_sum6(v, lds);
    if (threadIdx.x < 6) out[threadIdx.x] = lds[threadIdx.x * NEMATIC_THREADS];
}

size_t nematic_scratch_doubles() { return 6 * NEMATIC_BLOCKS + 6; }

hipError_t launch_nematic_sums(hipStream_t s, const Particles &p, int n, double *scratch) {
    static_assert(NEMATIC_BLOCKS == NEMATIC_THREADS, "the final pass reads one partial per thread");
    hipLaunchKernelGGL(k_nematic_partial, dim3(NEMATIC_BLOCKS), dim3(NEMATIC_THREADS), 0, s, p, n, scratch);
    hipLaunchKernelGGL(k_nematic_final, dim3(1), dim3(NEMATIC_THREADS), 0, s, scratch, scratch + 6 * NEMATIC_BLOCKS);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_fix_all(Particles p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double R[9];
    for (int c = 0; c < 9; c++) R[c] = p.o[c][i];
    fix_rotation(R);
    for (int c = 0; c < 9; c++) p.o[c][i] = R[c];
}

hipError_t launch_fix_all(hipStream_t s, const Particles &p, int n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fix_all, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------------------
// Packings without a neighbour grid: every pair through the nearest image (PeriodicBoundaryConditions::getTranslation
// for spherocylinders, getDistance2 for spheres; SpherocylinderTraits.cpp:28-46, SphereTraits.cpp:56-65)

constexpr int ALLPAIRS_THREADS = 256;         // overlap count: a thread per particle
constexpr int ALLPAIRS_SWEEP_THREADS = 1024;  // sweep: one workgroup, the candidates of a move spread over its lanes
constexpr int ALLPAIRS_CHUNK = 256;           // sweep: proposals staged in LDS at a time

template <int KIND>
__device__ __forceinline__ bool allpairs_hit(const PairConsts &k, const Grid &g, const SCFirst &f, const V3 &p2,
                                             const V3 &a2) {
    if (KIND == 0) return pbc_distance2(g.dims, g.inv, f.pos, p2) < k.contact2;
    return sc_overlap(k, f, p2, a2, pbc_translation(g.dims, g.inv, f.pos, p2));
}

constexpr int AP_GROUP = 4;  // moves on distinct particles tested at once by the grouped all-pairs sweep

size_t allpairs_lds_bytes(int n) {
    return sizeof(double) * 12 * (size_t)n + (sizeof(double) * 12 + sizeof(int)) * ALLPAIRS_CHUNK +
           sizeof(double) * 12 * AP_GROUP;
}

// The cycle's moves of a packing without a grid, in order, by one workgroup (Simulation::performMoves without domains,
// Simulation.cpp:365-445: every move picks any particle; Packing::tryMove / countParticleOverlaps without a grid,
// Packing.cpp:136-164, 410-420): the trial pose against every other particle, accepted poses written back at once.
// fixMode: the orientation renormalisation (Packing::renormalizeOrientations), one trial per particle in storage
// order, fix_rotation(O) in place.
// LDS: the packing (positions and the nine orientation components, P.cap particles at most) and the proposals,
// ALLPAIRS_CHUNK moves at a time, are staged in LDS and every move reads LDS only; accepted poses are written to
// both.  Otherwise every read goes to global memory (packings too large for LDS).
template <int KIND, bool LDS>
__global__ __launch_bounds__(ALLPAIRS_SWEEP_THREADS) void k_sweep_allpairs(SweepParams P) {
    extern __shared__ double smem[];
    constexpr int T = ALLPAIRS_SWEEP_THREADS, C = ALLPAIRS_CHUNK;
    const Grid &g = P.g;
    const int cell = 0;  // the one cell of a packing without a grid
    const int n = g.cellStart[1];
    const int tid = threadIdx.x;
    const int moves = P.fixMode ? n : P.moves;
    if (LDS && n > P.cap) {
        // the launch's LDS holds P.cap particles
        if (tid == 0) atomicAdd(&P.counters[CNT_STAGE_OVERFLOW], 1ull);
        return;
    }
    double *sx = nullptr, *sy = nullptr, *sz = nullptr, *so = nullptr, *ct = nullptr, *cR = nullptr;
    int *ck = nullptr;
    if (LDS) {
        sx = smem;
        sy = sx + n;
        sz = sy + n;
        so = sz + n;  // component c of particle j at so[c * n + j]
        ct = so + 9 * n;
        cR = ct + 3 * C;
        ck = (int *)(cR + 9 * C);
        for (int j = tid; j < n; j += T) {
            sx[j] = P.p.x[j];
            sy[j] = P.p.y[j];
            sz[j] = P.p.z[j];
            for (int c = 0; c < 9; c++) so[c * n + j] = P.p.o[c][j];
        }
    }
    auto pos_of = [&](int j) { return LDS ? v3(sx[j], sy[j], sz[j]) : load3(P.p.x, P.p.y, P.p.z, j); };
    auto o_of = [&](int c, int j) { return LDS ? so[c * n + j] : P.p.o[c][j]; };
    unsigned long long accepted = 0, tests = 0;
    for (int base = 0; base < moves; base += C) {
        const int cn = min(C, moves - base);
        if (LDS) {
            __syncthreads();  // the packing staged; every thread done with the previous chunk's proposals
            if (!P.fixMode && tid < cn) {
                const int i = base + tid;
                ck[tid] = P.pr.k[i];
                for (int d = 0; d < 3; d++) ct[d * C + tid] = P.pr.t[d * P.pr.stride + i];
                for (int c = 0; c < 9; c++) cR[c * C + tid] = P.pr.R[c * P.pr.stride + i];
            }
            __syncthreads();
        }
        for (int mm = 0; mm < cn; mm++) {
            const int m = base + mm;
            const int kIdx = P.fixMode ? m : (LDS ? ck[mm] : P.pr.k[m]);
            const V3 pos = pos_of(kIdx);
            V3 npos = pos;
            double O[9], R[9], NO[9];
            for (int c = 0; c < 9; c++) O[c] = o_of(c, kIdx);
            if (P.fixMode) {
                for (int c = 0; c < 9; c++) NO[c] = O[c];
                fix_rotation(NO);
            } else {
                const V3 t = LDS ? v3(ct[mm], ct[C + mm], ct[2 * C + mm])
                                 : v3(P.pr.t[m], P.pr.t[P.pr.stride + m], P.pr.t[2 * P.pr.stride + m]);
                npos = translate_wrapped(g.dims, g.inv, pos, t);
                for (int c = 0; c < 9; c++) R[c] = LDS ? cR[c * C + mm] : P.pr.R[c * P.pr.stride + m];
                mulm(R, O, NO);
            }
            const SCFirst f = make_first<KIND>(P.k, npos, v3(NO[2], NO[5], NO[8]));
            bool hit = false;
            for (int j = tid; j < n && !hit; j += T) {
                if (j == kIdx) continue;
                const V3 a2 = KIND == 1 ? v3(o_of(2, j), o_of(5, j), o_of(8, j)) : v3(0, 0, 0);
                hit = allpairs_hit<KIND>(P.k, g, f, pos_of(j), a2);
            }
            tests += (unsigned long long)(n - 1);
            if (__syncthreads_or(hit)) continue;
            accepted++;
            if (tid == 0) {
                P.p.x[kIdx] = npos.x;
                P.p.y[kIdx] = npos.y;
                P.p.z[kIdx] = npos.z;
                for (int c = 0; c < 9; c++) P.p.o[c][kIdx] = NO[c];
                P.p.wbits[kIdx] = 0;
                if (P.pr.acc && !P.fixMode) P.pr.acc[m] = 1;
                if (LDS) {
                    sx[kIdx] = npos.x;
                    sy[kIdx] = npos.y;
                    sz[kIdx] = npos.z;
                    for (int c = 0; c < 9; c++) so[c * n + kIdx] = NO[c];
                }
            }
            __syncthreads();
        }
    }
    flush_counters(P, cell, accepted, moves, 0, tests, true);
}

// Grouped form of the staged all-pairs sweep: up to AP_GROUP consecutive moves on distinct particles are tested at
// once against every particle and then decided in their order exactly as one at a time -- move k is rejected if it
// overlaps a particle outside the group, the old pose of a group particle i not accepted before it (i > k, or i < k
// rejected), or the trial pose of an accepted move i < k -- the rule of the grouped checkerboard sweep.  Same
// decisions and trajectory as k_sweep_allpairs<KIND, true>; three barriers per group instead of one or two per move.
template <int KIND>
__global__ __launch_bounds__(ALLPAIRS_SWEEP_THREADS) void k_sweep_allpairs_grouped(SweepParams P) {
    extern __shared__ double smem[];
    __shared__ unsigned sHit[2];  // [0]: 5 bits per move (bit i < 4: old pose of member i, bit 4: outside the group)
                                  // [1]: trial pairs, bit 4k + i: move k's trial pose overlaps move i's (i < k)
    constexpr int T = ALLPAIRS_SWEEP_THREADS, C = ALLPAIRS_CHUNK, G = AP_GROUP;
    const Grid &g = P.g;
    const int cell = 0;  // the one cell of a packing without a grid
    const int n = g.cellStart[1];
    const int tid = threadIdx.x;
    const int moves = P.moves;
    if (n > P.cap) {
        // the launch's LDS holds P.cap particles
        if (tid == 0) atomicAdd(&P.counters[CNT_STAGE_OVERFLOW], 1ull);
        return;
    }
    double *sx = smem, *sy = sx + n, *sz = sy + n;
    double *so = sz + n;  // component c of particle j at so[c * n + j]
    double *ct = so + 9 * n, *cR = ct + 3 * C;
    int *ck = (int *)(cR + 9 * C);
    double *plan = (double *)(ck + C);  // per group member: trial position (3) and orientation (9)
    for (int j = tid; j < n; j += T) {
        sx[j] = P.p.x[j];
        sy[j] = P.p.y[j];
        sz[j] = P.p.z[j];
        for (int c = 0; c < 9; c++) so[c * n + j] = P.p.o[c][j];
    }
    unsigned long long accepted = 0, tests = 0;
    for (int base = 0; base < moves; base += C) {
        const int cn = min(C, moves - base);
        __syncthreads();  // the packing staged; every thread done with the previous chunk's proposals
        if (tid < cn) {
            const int i = base + tid;
            ck[tid] = P.pr.k[i];
            for (int d = 0; d < 3; d++) ct[d * C + tid] = P.pr.t[d * P.pr.stride + i];
            for (int c = 0; c < 9; c++) cR[c * C + tid] = P.pr.R[c * P.pr.stride + i];
        }
        __syncthreads();
        for (int m0 = 0; m0 < cn;) {
            // the group: the longest run of at most G moves on distinct particles (the same in every lane)
            const int c0 = ck[m0];
            const int c1 = m0 + 1 < cn ? ck[m0 + 1] : -1, c2 = m0 + 2 < cn ? ck[m0 + 2] : -1;
            const int c3 = m0 + 3 < cn ? ck[m0 + 3] : -1;
            const bool d1 = m0 + 1 < cn && c1 != c0;
            const bool d2 = d1 && m0 + 2 < cn && c2 != c0 && c2 != c1;
            const bool d3 = d2 && m0 + 3 < cn && c3 != c0 && c3 != c1 && c3 != c2;
            const int gs = 1 + (int)d1 + (int)d2 + (int)d3;
            const int k0 = c0, k1 = d1 ? c1 : -1, k2 = d2 ? c2 : -1, k3 = d3 ? c3 : -1;
            auto kof = [&](int k) { return k == 0 ? k0 : (k == 1 ? k1 : (k == 2 ? k2 : k3)); };
            if (tid < gs) {
                // trial pose of move m0 + tid (Shape::translate with the periodic correction, Shape::rotate)
                const int mm = m0 + tid, ki = kof(tid);
                const V3 npos = translate_wrapped(g.dims, g.inv, v3(sx[ki], sy[ki], sz[ki]),
                                                  v3(ct[mm], ct[C + mm], ct[2 * C + mm]));
                double O[9], R[9], NO[9];
                for (int c = 0; c < 9; c++) {
                    O[c] = so[c * n + ki];
                    R[c] = cR[c * C + mm];
                }
                mulm(R, O, NO);
                double *q = plan + 12 * tid;
                q[0] = npos.x;
                q[1] = npos.y;
                q[2] = npos.z;
                for (int c = 0; c < 9; c++) q[3 + c] = NO[c];
            }
            if (tid < 2) sHit[tid] = 0;
            __syncthreads();
            unsigned hw = 0, tw = 0;
            for (int k = 0; k < gs; k++) {
                const double *q = plan + 12 * k;
                const SCFirst f = make_first<KIND>(P.k, v3(q[0], q[1], q[2]), v3(q[5], q[8], q[11]));
                const int kk = kof(k);
                for (int j = tid; j < n; j += T) {
                    if (j == kk) continue;
                    const int member = j == k0 ? 0 : (j == k1 ? 1 : (j == k2 ? 2 : (j == k3 ? 3 : 4)));
                    const unsigned bit = 1u << (5 * k + member);
                    if (hw & bit) continue;
                    const V3 a2 = KIND == 1 ? v3(so[2 * n + j], so[5 * n + j], so[8 * n + j]) : v3(0, 0, 0);
                    if (allpairs_hit<KIND>(P.k, g, f, v3(sx[j], sy[j], sz[j]), a2)) hw |= bit;
                }
                // lanes 4k + i test move k's trial pose against move i's (i < k)
                for (int i = 0; i < k; i++) {
                    if (tid != 4 * k + i) continue;
                    const double *qi = plan + 12 * i;
                    const V3 a2 = KIND == 1 ? v3(qi[5], qi[8], qi[11]) : v3(0, 0, 0);
                    if (allpairs_hit<KIND>(P.k, g, f, v3(qi[0], qi[1], qi[2]), a2)) tw |= 1u << (4 * k + i);
                }
            }
            if (hw) atomicOr(&sHit[0], hw);
            if (tw) atomicOr(&sHit[1], tw);
            __syncthreads();
            // the group's decisions in order (the same in every lane)
            const unsigned H = sHit[0], TP = sHit[1];
            unsigned acc = 0;
            for (int k = 0; k < gs; k++) {
                const unsigned b = (H >> (5 * k)) & 31u, trial = (TP >> (4 * k)) & 15u;
                const bool hit = (b & 16u) || ((b & 15u) & ~acc) || (trial & acc);
                if (!hit) acc |= 1u << k;
            }
            accepted += (unsigned long long)__popc(acc);
            tests += (unsigned long long)(n - 1) * (unsigned long long)gs;
            if (tid < gs && ((acc >> tid) & 1u)) {
                const int ki = kof(tid);
                const double *q = plan + 12 * tid;
                P.p.x[ki] = q[0];
                P.p.y[ki] = q[1];
                P.p.z[ki] = q[2];
                for (int c = 0; c < 9; c++) P.p.o[c][ki] = q[3 + c];
                P.p.wbits[ki] = 0;
                if (P.pr.acc) P.pr.acc[base + m0 + tid] = 1;
                sx[ki] = q[0];
                sy[ki] = q[1];
                sz[ki] = q[2];
                for (int c = 0; c < 9; c++) so[c * n + ki] = q[3 + c];
            }
            __syncthreads();  // accepted poses staged; the plan and hit words free for the next group
            m0 += gs;
        }
    }
    flush_counters(P, cell, accepted, moves, 0, tests, true);
}

template <int KIND>
__global__ __launch_bounds__(ALLPAIRS_THREADS) void k_count_allpairs(Grid g, int n, Particles p, PairConsts k,
                                                                    int earlyExit, unsigned long long *counters) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int ti = p.tag[i];
    const SCFirst f = make_first<KIND>(k, load3(p.x, p.y, p.z, i), v3(p.o[2][i], p.o[5][i], p.o[8][i]));
    unsigned long long c = 0;
    for (int j = 0; j < n; j++) {
        if (p.tag[j] <= ti) continue;  // each pair once, from the particle with the lower caller index
        const V3 a2 = KIND == 1 ? v3(p.o[2][j], p.o[5][j], p.o[8][j]) : v3(0, 0, 0);
        if (allpairs_hit<KIND>(k, g, f, load3(p.x, p.y, p.z, j), a2)) {
            c++;
            if (earlyExit) break;
        }
    }
    if (c) atomicAdd(&counters[CNT_OVERLAPS], c);
}

hipError_t launch_count_allpairs(hipStream_t s, const Grid &g, int n, Particles p, PairConsts k, int earlyExit,
                                 unsigned long long *counters) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + ALLPAIRS_THREADS - 1) / ALLPAIRS_THREADS);
    if (k.kind == 0)
        hipLaunchKernelGGL(k_count_allpairs<0>, dim3(blocks), dim3(ALLPAIRS_THREADS), 0, s, g, n, p, k, earlyExit,
                           counters);
    else if (k.kind == 1)
        hipLaunchKernelGGL(k_count_allpairs<1>, dim3(blocks), dim3(ALLPAIRS_THREADS), 0, s, g, n, p, k, earlyExit,
                           counters);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

int sweep_occupancy(int kind, int cap, int waves) {
    int w = waves >= 8 ? 8 : (waves >= 4 ? 4 : (waves == 2 ? 2 : 1));
    size_t lds = sweep_lds_bytes(kind, cap, w);
    int blocks = 0;
#define RP_OCC(KIND, W) hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_sweep<KIND, W, false>, WAVE * W, lds)
    if (kind == 0) {
        if (w == 8) RP_OCC(0, 8);
        else if (w == 4) RP_OCC(0, 4);
        else if (w == 2) RP_OCC(0, 2);
        else RP_OCC(0, 1);
    } else {
        if (w == 8) RP_OCC(1, 8);
        else if (w == 4) RP_OCC(1, 4);
        else if (w == 2) RP_OCC(1, 2);
        else RP_OCC(1, 1);
    }
#undef RP_OCC
    return blocks;
}

hipError_t launch_sweep(hipStream_t s, const SweepParams &sp, int waves, Timed t, int g32) {
    if (sp.g.allPairs) {
        // no grid: one workgroup runs the cycle's moves in order (hard spheres and spherocylinders)
        if (g32 || sp.k.soft || sp.counting || sp.k.kind == 2 || (!sp.fixMode && sp.moves <= 0))
            return hipErrorInvalidValue;
        // sp.cap > 0: the packing (at most sp.cap particles) and the proposals staged in LDS
        const uint32_t lds = sp.cap > 0 ? (uint32_t)allpairs_lds_bytes(sp.cap) : 0;
#define RP_ALLPAIRS(KIND, STAGED) \
    hipExtLaunchKernelGGL((k_sweep_allpairs<KIND, STAGED>), dim3(1), dim3(ALLPAIRS_SWEEP_THREADS), lds, s, t.start, t.stop, \
                          0, sp)
#define RP_ALLPAIRS_GROUPED(KIND) \
    hipExtLaunchKernelGGL((k_sweep_allpairs_grouped<KIND>), dim3(1), dim3(ALLPAIRS_SWEEP_THREADS), lds, s, t.start, \
                          t.stop, 0, sp)
        // staged packings: grouped moves; the orientation fix and packings too large for LDS: one move at a time
        const bool grouped = sp.cap > 0 && !sp.fixMode;
        if (sp.k.kind == 0 && grouped) RP_ALLPAIRS_GROUPED(0);
        else if (sp.k.kind == 0 && sp.cap > 0) RP_ALLPAIRS(0, true);
        else if (sp.k.kind == 0) RP_ALLPAIRS(0, false);
        else if (grouped) RP_ALLPAIRS_GROUPED(1);
        else if (sp.cap > 0) RP_ALLPAIRS(1, true);
        else RP_ALLPAIRS(1, false);
#undef RP_ALLPAIRS_GROUPED
#undef RP_ALLPAIRS
        return hipGetLastError();
    }
    unsigned blocks = (unsigned)(sp.na[0] * sp.na[1] * sp.na[2]);
    int w = waves >= 16 ? 16 : (waves >= 8 ? 8 : (waves >= 4 ? 4 : (waves == 2 ? 2 : 1)));
    if (w == 16 && (sp.k.kind != 2 || sp.counting || sp.fixMode)) w = 8;  // sixteen only for the polysphere sweep
    if (w == 8 && (sp.k.soft || sp.counting)) w = 4;  // eight only for the hard sweeps
    // the lean kernels run spherocylinders only: their first launch on spheres (eight wavefronts forced onto sparse
    // stencils) faulted the GPU and the cause is not found, so spheres keep the general kernel
    if (g32 && (w != 8 || sp.pipelined != 2 || sp.fixMode || sp.k.kind != 1)) return hipErrorInvalidValue;
    size_t lds = sweep_lds_bytes(sp.k.kind, sp.cap, w, g32);
    // the default hard launch (eight wavefronts, grouped moves) runs the lean kernels
    const bool lean = w == 8 && sp.pipelined == 2 && !sp.fixMode && !sp.k.soft && !sp.counting && sp.k.kind == 1;
    if (lean) {
#define RP_SWEEP_L(L) \
    hipExtLaunchKernelGGL((k_sweep<1, 8, false, L>), dim3(blocks), dim3(WAVE * 8), (uint32_t)lds, s, t.start, t.stop, 0, sp)
        const int mode = 1 + g32;
        if (mode == 3) RP_SWEEP_L(3);
        else if (mode == 2) RP_SWEEP_L(2);
        else RP_SWEEP_L(1);
#undef RP_SWEEP_L
        return hipGetLastError();
    }
    // soft and hard sweeps are separate kernels: each instantiation carries one path only
#define RP_SWEEP(KIND, W, SOFT) \
    hipExtLaunchKernelGGL((k_sweep<KIND, W, SOFT>), dim3(blocks), dim3(WAVE * W), (uint32_t)lds, s, t.start, t.stop, 0, sp)
#define RP_SWEEP_W(KIND, SOFT)         \
    if (w == 4) RP_SWEEP(KIND, 4, SOFT); \
    else if (w == 2) RP_SWEEP(KIND, 2, SOFT); \
    else RP_SWEEP(KIND, 1, SOFT);
    if (sp.k.kind == 2) {
        // polyspheres: the hard sweep reading the stencil from HBM; counting mode through the soft sweep's counting
        // branch (at most four wavefronts)
        if (sp.counting) { RP_SWEEP_W(2, true) }
        else if (w == 16) RP_SWEEP(2, 16, false);
        else if (w == 8) RP_SWEEP(2, 8, false);
        else { RP_SWEEP_W(2, false) }
    } else if (sp.k.kind == 0) {
        if (sp.k.soft || sp.counting) { RP_SWEEP_W(0, true) }
        else if (w == 8) RP_SWEEP(0, 8, false);
        else { RP_SWEEP_W(0, false) }
    } else {
        if (sp.k.soft || sp.counting) { RP_SWEEP_W(1, true) }
        else if (w == 8) RP_SWEEP(1, 8, false);
        else { RP_SWEEP_W(1, false) }
    }
#undef RP_SWEEP_W
#undef RP_SWEEP
    return hipGetLastError();
}

int sweep_batch_key(const SweepParams &sp, int waves, int g32) {
    if (sp.k.soft || sp.counting || sp.fixMode || sp.k.kind == 2 || sp.trace || sp.g.allPairs) return -1;
    int w = waves >= 8 ? 8 : (waves >= 4 ? 4 : (waves == 2 ? 2 : 1));
    const bool lean = w == 8 && sp.pipelined == 2 && sp.k.kind == 1;  // as launch_sweep chooses
    if (g32 && !lean) return -1;
    return sp.k.kind | (w << 4) | ((lean ? 1 + g32 : 0) << 8);
}

hipError_t launch_sweep_batch(hipStream_t s, int key, const SweepParams *table, const int *first, int packings,
                              unsigned blocks, size_t lds) {
    if (blocks == 0) return hipSuccess;
    const int kind = key & 15, w = (key >> 4) & 15, lean = key >> 8;
#define RP_BATCH(KIND, W, L) \
    hipLaunchKernelGGL((k_sweep_batch<KIND, W, L>), dim3(blocks), dim3(WAVE * W), (uint32_t)lds, s, table, first, packings)
    if (kind == 1 && lean) {
        if (lean == 3) RP_BATCH(1, 8, 3);
        else if (lean == 2) RP_BATCH(1, 8, 2);
        else RP_BATCH(1, 8, 1);
    } else if (kind == 0) {
        if (w == 8) RP_BATCH(0, 8, 0);
        else if (w == 4) RP_BATCH(0, 4, 0);
        else if (w == 2) RP_BATCH(0, 2, 0);
        else RP_BATCH(0, 1, 0);
    } else if (kind == 1) {
        if (w == 8) RP_BATCH(1, 8, 0);
        else if (w == 4) RP_BATCH(1, 4, 0);
        else if (w == 2) RP_BATCH(1, 2, 0);
        else RP_BATCH(1, 1, 0);
    } else {
        return hipErrorInvalidValue;
    }
#undef RP_BATCH
    return hipGetLastError();
}

}  // namespace rp
