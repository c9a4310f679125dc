// rp_kernels.hip -- hand-written gfx950 kernels for the rampack MC sweep.
//
//   * k_pair_decisions   batched FP64 pair-overlap decisions (the decision function of
//                        SpherocylinderTraits::overlapBetween / SphereTraits, SpherocylinderTraits.cpp:28-46,
//                        SphereTraits.cpp:56-65) on explicit (pose, pose, translation) records
//   * k_bin / k_scan / k_scatter / k_gather
//                        counting-sort cell-list rebuild over coalesced SoA records, replacing the linked-list
//                        NeighbourGrid insertion (Packing.cpp:858-929, NeighbourGrid.cpp:219-227); the order
//                        inside a cell is the stable order of the source arrays, so rebuilds are deterministic
//   * k_count_overlaps   half-shell overlap count with early exit (Packing.cpp:431-559)
//   * k_sweep            checkerboard MC sweep: one wavefront per active cell, the 27-cell stencil staged in
//                        LDS, early-exit overlap voting by wavefront ballot, a Philox stream per cell, moves
//                        that leave the cell rejected (the ActiveDomain rule of Packing.cpp:147-148)
#include "rp_kernels.h"

#include <cfloat>

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
            cc = (int)(y / g.relCell[d]);
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

__device__ __forceinline__ V3 translation(const Grid &g, int kx, int ky, int kz) {
    int t = 3 * ((kx + 1) * 9 + (ky + 1) * 3 + (kz + 1));
    return v3(g.T[t], g.T[t + 1], g.T[t + 2]);
}

template <int KIND>
__device__ __forceinline__ bool pair_overlap(const PairConsts &k, const SCFirst &f, const V3 &p2, const V3 &a2,
                                             const V3 &T) {
    if (KIND == 0) return sphere_overlap(k, f.pos, p2, T);
    return sc_overlap(k, f, p2, a2, T);
}

template <int KIND>
__device__ __forceinline__ SCFirst make_first(const PairConsts &k, const V3 &pos, const V3 &axis) {
    if (KIND == 0) {
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
    bool o;
    if (k.kind == 0) {
        o = sphere_overlap(k, p1, p2, T);
    } else {
        SCFirst f = sc_first(k, p1, a1);
        o = sc_overlap(k, f, p2, a2, T);
    }
    out[i] = o ? 1 : 0;
}

hipError_t launch_pair_decisions(hipStream_t s, PairConsts k, size_t n, const double *rec, uint8_t *out) {
    if (n == 0) return hipSuccess;
    unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_pair_decisions, dim3(blocks), dim3(256), 0, s, k, n, rec, out);
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
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3 pos = load3(src.x, src.y, src.z, i);
    if (rescale) pos = mulv(g.dims, mulv(oldInv.m, pos));
    tmpPos[i] = pos.x;
    tmpPos[n + i] = pos.y;
    tmpPos[2 * n + i] = pos.z;
    int c[3];
    unsigned w;
    bin_position(g, pos, c, w);
    int cell = cell_index(g, c[0], c[1], c[2]);
    tmpCell[i] = cell;
    tmpW[i] = (uint8_t)w;
    tmpSlot[i] = atomicAdd(&g.cellCount[cell], 1);
}

// Exclusive scan of the cell counts by one workgroup: contiguous chunk per thread, then a scan of the chunk sums.
__global__ __launch_bounds__(1024) void k_scan(int ncell, const int *__restrict__ count, int *__restrict__ start,
                                               unsigned long long *counters) {
    __shared__ int part[1024];
    __shared__ int pmax[1024];
    int t = threadIdx.x;
    int per = (ncell + 1023) / 1024;
    int b = t * per, e = min(b + per, ncell);
    int s = 0, mx = 0;
    for (int i = b; i < e; i++) {
        int c = count[i];
        s += c;
        mx = max(mx, c);
    }
    part[t] = s;
    pmax[t] = mx;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = t >= off ? part[t - off] : 0;
        int m = t >= off ? pmax[t - off] : 0;
        __syncthreads();
        part[t] += v;
        pmax[t] = max(pmax[t], m);
        __syncthreads();
    }
    int run = t == 0 ? 0 : part[t - 1];
    for (int i = b; i < e; i++) {
        start[i] = run;
        run += count[i];
    }
    if (t == 1023) {
        start[ncell] = part[1023];
        if (counters) atomicMax(&counters[CNT_MAX_STENCIL], (unsigned long long)pmax[1023]);
    }
}

__global__ __launch_bounds__(256) void k_scatter(int n, const int *__restrict__ cellStart,
                                                 const int *__restrict__ tmpCell, const int *__restrict__ tmpSlot,
                                                 int *__restrict__ tmpOrder) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tmpOrder[cellStart[tmpCell[i]] + tmpSlot[i]] = i;
}

// One wavefront per cell: rank the cell's members by source index (a stable order), then gather their records
// into the destination arrays.
__global__ __launch_bounds__(WAVE) void k_gather(Grid g, int n, Particles src, Particles dst,
                                                 const double *__restrict__ tmpPos,
                                                 const uint8_t *__restrict__ tmpW,
                                                 const int *__restrict__ tmpOrder) {
    int cell = blockIdx.x;
    int lane = threadIdx.x;
    int start = g.cellStart[cell];
    int m = g.cellStart[cell + 1] - start;
    if (m == 0) return;
    for (int base = 0; base < m; base += WAVE) {
        int e = base + lane;
        int srcIdx = e < m ? tmpOrder[start + e] : 0x7fffffff;
        int rank = 0;
        if (m <= WAVE) {
            for (int q = 0; q < m; q++) {
                int other = __shfl(srcIdx, q, WAVE);
                rank += other < srcIdx ? 1 : 0;
            }
        } else if (e < m) {
            for (int q = 0; q < m; q++) rank += tmpOrder[start + q] < srcIdx ? 1 : 0;
        }
        if (e < m) {
            int d = start + rank;
            dst.x[d] = tmpPos[srcIdx];
            dst.y[d] = tmpPos[n + srcIdx];
            dst.z[d] = tmpPos[2 * n + srcIdx];
            for (int c = 0; c < 9; c++) dst.o[c][d] = src.o[c][srcIdx];
            dst.tag[d] = src.tag[srcIdx];
            dst.wbits[d] = tmpW[srcIdx];
            dst.cell[d] = cell;
        }
    }
}

hipError_t launch_rebuild(hipStream_t s, const Grid &g, int n, Particles src, Particles dst, int *tmpSlot,
                          int *tmpOrder, double *tmpPos, int rescale, const double *oldInv) {
    // tmpSlot holds 2n ints (slot, cell); tmpOrder holds n ints followed by n bytes of wrap bits.
    int *tmpCell = tmpSlot + n;
    uint8_t *tmpW = reinterpret_cast<uint8_t *>(tmpOrder + n);
    hipError_t err = launch_fill_i32(s, g.cellCount, (size_t)g.ncell, 0);
    if (err != hipSuccess) return err;
    M3 inv{};
    if (rescale)
        for (int i = 0; i < 9; i++) inv.m[i] = oldInv[i];
    unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_bin, dim3(blocks), dim3(256), 0, s, g, n, src, rescale, inv, tmpPos, tmpCell, tmpSlot, tmpW);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, g.ncell, g.cellCount, g.cellStart,
                       (unsigned long long *)nullptr);
    hipLaunchKernelGGL(k_scatter, dim3(blocks), dim3(256), 0, s, n, g.cellStart, tmpCell, tmpSlot, tmpOrder);
    hipLaunchKernelGGL(k_gather, dim3((unsigned)g.ncell), dim3(WAVE), 0, s, g, n, src, dst, tmpPos, tmpW, tmpOrder);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------------------
// Half-shell overlap count (Packing::countTotalOverlapsNGCellHelper, Packing.cpp:431-517).  Within a cell the
// particle with the higher caller index is the first operand, as in the reference's LIFO cell lists; across cells
// the first operand lies in the base cell and the second in one of the 13 "positive" neighbours
// (9*n0 + 3*n1 + n2 > 13, NeighbourGrid.cpp:130-131).

template <int KIND>
__global__ __launch_bounds__(256) void k_count_overlaps(Grid g, int n, Particles p, PairConsts k, int earlyExit,
                                                        unsigned long long *counters) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    volatile unsigned long long *flag = counters + CNT_OVERLAPS;
    if (earlyExit && *flag != 0) return;
    int cell = p.cell[i];
    int cx = cell % g.n[0], cy = (cell / g.n[0]) % g.n[1], cz = cell / (g.n[0] * g.n[1]);
    int wi = p.wbits[i];
    int tagi = p.tag[i];
    V3 pos1 = load3(p.x, p.y, p.z, i);
    V3 ax1 = KIND == 1 ? v3(p.o[2][i], p.o[5][i], p.o[8][i]) : v3(0, 0, 0);
    SCFirst f = make_first<KIND>(k, pos1, ax1);
    unsigned long long found = 0;
    // own cell
    {
        int b = g.cellStart[cell], e = g.cellStart[cell + 1];
        for (int j = b; j < e; j++) {
            if (p.tag[j] >= tagi) continue;
            int wj = p.wbits[j];
            V3 T = v3(0, 0, 0);
            if (wj != wi)
                T = translation(g, (wj & 1) - (wi & 1), ((wj >> 1) & 1) - ((wi >> 1) & 1),
                                ((wj >> 2) & 1) - ((wi >> 2) & 1));
            V3 p2 = load3(p.x, p.y, p.z, j);
            V3 a2 = KIND == 1 ? v3(p.o[2][j], p.o[5][j], p.o[8][j]) : v3(0, 0, 0);
            if (pair_overlap<KIND>(k, f, p2, a2, T)) {
                found++;
                if (earlyExit) break;
            }
        }
    }
    for (int q = 14; q < 27 && !(earlyExit && found); q++) {
        if (earlyExit && *flag != 0) break;
        int ex, ey, ez;
        int nx = wrap_axis(cx, q / 9 - 1, g.n[0], ex);
        int ny = wrap_axis(cy, (q / 3) % 3 - 1, g.n[1], ey);
        int nz = wrap_axis(cz, q % 3 - 1, g.n[2], ez);
        int nc = cell_index(g, nx, ny, nz);
        int b = g.cellStart[nc], e = g.cellStart[nc + 1];
        for (int j = b; j < e; j++) {
            int wj = p.wbits[j];
            V3 T = translation(g, ex + (wj & 1) - (wi & 1), ey + ((wj >> 1) & 1) - ((wi >> 1) & 1),
                               ez + ((wj >> 2) & 1) - ((wi >> 2) & 1));
            V3 p2 = load3(p.x, p.y, p.z, j);
            V3 a2 = KIND == 1 ? v3(p.o[2][j], p.o[5][j], p.o[8][j]) : v3(0, 0, 0);
            if (pair_overlap<KIND>(k, f, p2, a2, T)) {
                found++;
                if (earlyExit) break;
            }
        }
    }
    if (found) atomicAdd(&counters[CNT_OVERLAPS], found);
}

hipError_t launch_count_overlaps(hipStream_t s, const Grid &g, int n, Particles p, PairConsts k, int earlyExit,
                                 unsigned long long *counters) {
    if (n == 0) return hipSuccess;
    unsigned blocks = (unsigned)((n + 255) / 256);
    if (k.kind == 0)
        hipLaunchKernelGGL(k_count_overlaps<0>, dim3(blocks), dim3(256), 0, s, g, n, p, k, earlyExit, counters);
    else
        hipLaunchKernelGGL(k_count_overlaps<1>, dim3(blocks), dim3(256), 0, s, g, n, p, k, earlyExit, counters);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------------------
// Checkerboard sweep.  Block = one wavefront = one active cell of colour `colour` (cells with coordinates of the
// colour's parity on every axis; same-colour cells are separated by a full inactive cell, wider than the
// interaction range).  The cell's particles are moved n_cell times (uniform random pick, rototranslation as
// RototranslationSampler.cpp:22-49), every candidate of the 27-cell stencil is tested from LDS, 64 candidates
// per step, and the first step with any overlap rejects the move.

size_t sweep_lds_bytes(int kind, int cap) {
    size_t perEntry = (kind == 0 ? 3 : 6) * sizeof(double) + sizeof(int);
    return perEntry * (size_t)cap;
}

// Stencil view of one block: entry j of the 27 stencil cells (own cell first) either staged in LDS or read from
// the cell-sorted arrays in HBM when the stencil does not fit the LDS area.
struct Stencil {
    const int *start;   // 28 prefix offsets (LDS)
    const int *base;    // 27 first record indices (LDS)
    const int *ecode;   // 27 packed cell image shifts (LDS)
    double *x, *y, *z, *ax, *ay, *az;
    int *code;
};

__device__ __forceinline__ int stencil_cell(const int *start, int j) {
    int lo = 0, hi = 27;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (start[mid] <= j) lo = mid; else hi = mid;
    }
    return lo;
}

// Per-axis image code e + w_j + 1 (two bits per axis) of a record with wrap bits w in a stencil cell with packed
// image shift ec.
__device__ __forceinline__ int image_code(int ec, int w) {
    return (((ec & 3) - 1 + (w & 1)) + 1) | ((((ec >> 2) & 3) - 1 + ((w >> 1) & 1) + 1) << 2) |
           ((((ec >> 4) & 3) - 1 + ((w >> 2) & 1) + 1) << 4);
}

template <int KIND, bool STAGED>
__device__ __forceinline__ void load_candidate(const SweepParams &P, const Stencil &S, int j, V3 &p2, V3 &a2,
                                               int &code) {
    if (STAGED) {
        p2 = v3(S.x[j], S.y[j], S.z[j]);
        a2 = KIND == 1 ? v3(S.ax[j], S.ay[j], S.az[j]) : v3(0, 0, 0);
        code = S.code[j];
    } else {
        int q = stencil_cell(S.start, j);
        int gi = S.base[q] + (j - S.start[q]);
        p2 = load3(P.p.x, P.p.y, P.p.z, gi);
        a2 = KIND == 1 ? v3(P.p.o[2][gi], P.p.o[5][gi], P.p.o[8][gi]) : v3(0, 0, 0);
        code = image_code(S.ecode[q], P.p.wbits[gi]);
    }
}

// The trial moves of one active cell: n_cell moves, each a uniform pick among the cell's particles and a
// rototranslation (RototranslationSampler.cpp:22-49), rejected if the particle leaves the cell, otherwise tested
// against the stencil 64 candidates at a time with a wavefront vote after every step.
template <int KIND, bool STAGED>
__device__ void sweep_cell(const SweepParams &P, const Stencil &S, int cell, int own0, int nOwn, int total) {
    const Grid &g = P.g;
    const int lane = threadIdx.x;
    Philox rng = philox_init(P.seed, (P.cycle << 3) | (uint32_t)P.colour, (uint32_t)cell, 0x5eed5u);
    unsigned long long accepted = 0, leftCell = 0, tests = 0;
    const double ts = P.transStep, rs = P.rotStep;

    for (int m = 0; m < nOwn; m++) {
        int kIdx = (int)philox_below(rng, (uint32_t)nOwn);
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
        double angle = philox_uniform(rng, -rs, rs);

        const int gi = own0 + kIdx;
        V3 pos = STAGED ? v3(S.x[kIdx], S.y[kIdx], S.z[kIdx]) : load3(P.p.x, P.p.y, P.p.z, gi);
        V3 npos = translate_wrapped(g.dims, g.inv, pos, t);
        int nc[3];
        unsigned nw;
        bin_position(g, npos, nc, nw);
        if (cell_index(g, nc[0], nc[1], nc[2]) != cell) {
            leftCell++;
            continue;
        }

        double O[9], R[9], NO[9];
        for (int c = 0; c < 9; c++) O[c] = P.p.o[c][gi];
        // Matrix<3,3>::rotation(axis.normalized(), angle) (Matrix.tpp:438-450), then Shape::rotate: R * O
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
        mulm(R, O, NO);
        V3 nax = v3(NO[2], NO[5], NO[8]);
        SCFirst f = make_first<KIND>(P.k, npos, nax);
        const int wx = nw & 1, wy = (nw >> 1) & 1, wz = (nw >> 2) & 1;

        bool hit = false;
        for (int base = 0; base < total; base += WAVE) {
            int j = base + lane;
            bool h = false;
            if (j < total && j != kIdx) {
                V3 p2, a2;
                int code;
                load_candidate<KIND, STAGED>(P, S, j, p2, a2, code);
                V3 T = translation(g, (code & 3) - 1 - wx, ((code >> 2) & 3) - 1 - wy, ((code >> 4) & 3) - 1 - wz);
                h = pair_overlap<KIND>(P.k, f, p2, a2, T);
            }
            tests += (unsigned long long)min(WAVE, total - base);
            if (__any(h)) {
                hit = true;
                break;
            }
        }
        if (hit) continue;

        // Packing::acceptMove: the cell belongs to this wavefront for the whole colour pass
        accepted++;
        if (lane == 0) {
            P.p.x[gi] = npos.x;
            P.p.y[gi] = npos.y;
            P.p.z[gi] = npos.z;
            for (int c = 0; c < 9; c++) P.p.o[c][gi] = NO[c];
            P.p.wbits[gi] = (uint8_t)nw;
            if (STAGED) {
                S.x[kIdx] = npos.x;
                S.y[kIdx] = npos.y;
                S.z[kIdx] = npos.z;
                if (KIND == 1) {
                    S.ax[kIdx] = nax.x;
                    S.ay[kIdx] = nax.y;
                    S.az[kIdx] = nax.z;
                }
                S.code[kIdx] = (wx + 1) | ((wy + 1) << 2) | ((wz + 1) << 4);
            }
        }
        __syncthreads();
    }
    if (lane == 0) {
        atomicAdd(&P.counters[CNT_ACCEPTED], accepted);
        atomicAdd(&P.counters[CNT_ATTEMPTED], (unsigned long long)nOwn);
        atomicAdd(&P.counters[CNT_LEFT_CELL], leftCell);
        atomicAdd(&P.counters[CNT_TESTS], tests);
        if (!STAGED) atomicAdd(&P.counters[CNT_UNSTAGED_BLOCKS], 1ull);
    }
}

template <int KIND>
__global__ __launch_bounds__(WAVE) void k_sweep(SweepParams P) {
    extern __shared__ double smem[];
    __shared__ int s_start[28];
    __shared__ int s_base[27];
    __shared__ int s_ecode[27];

    const Grid &g = P.g;
    const int lane = threadIdx.x;
    const int b = blockIdx.x;
    const int bx = b % P.na[0], by = (b / P.na[0]) % P.na[1], bz = b / (P.na[0] * P.na[1]);
    const int cx = 2 * bx + (P.colour & 1), cy = 2 * by + ((P.colour >> 1) & 1), cz = 2 * bz + ((P.colour >> 2) & 1);
    const int cell = cell_index(g, cx, cy, cz);
    const int own0 = g.cellStart[cell];
    const int nOwn = g.cellStart[cell + 1] - own0;
    if (nOwn == 0) return;

    // Stencil cells: index 0 is the own cell, then the 26 neighbours.
    if (lane < 27) {
        int t = lane == 0 ? 13 : (lane <= 13 ? lane - 1 : lane);
        int ex, ey, ez;
        int nx = wrap_axis(cx, t % 3 - 1, g.n[0], ex);
        int ny = wrap_axis(cy, (t / 3) % 3 - 1, g.n[1], ey);
        int nz = wrap_axis(cz, t / 9 - 1, g.n[2], ez);
        int nc = cell_index(g, nx, ny, nz);
        s_base[lane] = g.cellStart[nc];
        s_start[lane + 1] = g.cellStart[nc + 1] - g.cellStart[nc];
        s_ecode[lane] = (ex + 1) | ((ey + 1) << 2) | ((ez + 1) << 4);
    }
    __syncthreads();
    if (lane == 0) {
        s_start[0] = 0;
        for (int q = 1; q <= 27; q++) s_start[q] += s_start[q - 1];
    }
    __syncthreads();
    const int total = s_start[27];

    Stencil S;
    S.start = s_start;
    S.base = s_base;
    S.ecode = s_ecode;
    S.x = smem;
    S.y = S.x + P.cap;
    S.z = S.y + P.cap;
    S.ax = S.z + P.cap;
    S.ay = S.ax + P.cap;
    S.az = S.ay + P.cap;
    S.code = reinterpret_cast<int *>(KIND == 0 ? S.ax : S.az + P.cap);

    if (total > P.cap) {
        sweep_cell<KIND, false>(P, S, cell, own0, nOwn, total);
        return;
    }
    for (int j = lane; j < total; j += WAVE) {
        int q = stencil_cell(s_start, j);
        int gi = s_base[q] + (j - s_start[q]);
        S.x[j] = P.p.x[gi];
        S.y[j] = P.p.y[gi];
        S.z[j] = P.p.z[gi];
        if (KIND == 1) {
            S.ax[j] = P.p.o[2][gi];
            S.ay[j] = P.p.o[5][gi];
            S.az[j] = P.p.o[8][gi];
        }
        S.code[j] = image_code(s_ecode[q], P.p.wbits[gi]);
    }
    __syncthreads();
    sweep_cell<KIND, true>(P, S, cell, own0, nOwn, total);
}

hipError_t launch_sweep(hipStream_t s, const SweepParams &sp) {
    unsigned blocks = (unsigned)(sp.na[0] * sp.na[1] * sp.na[2]);
    size_t lds = sweep_lds_bytes(sp.k.kind, sp.cap);
    if (sp.k.kind == 0)
        hipLaunchKernelGGL(k_sweep<0>, dim3(blocks), dim3(WAVE), lds, s, sp);
    else
        hipLaunchKernelGGL(k_sweep<1>, dim3(blocks), dim3(WAVE), lds, s, sp);
    return hipGetLastError();
}

}  // namespace rp
