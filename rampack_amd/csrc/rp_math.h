// rp_math.h -- FP64 geometry for the MC sweep, host and device.
//
// Every function evaluates in the operation order of the rampack routine it
// replaces (cited path:line into the reference tree), with each * and + rounded
// on its own: the whole library is compiled with -ffp-contract=off, so a pair
// decision computed here is bit-identical to the reference's FP64 decision on
// the same operands.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#define RP_HD __host__ __device__ __forceinline__

namespace rp {

struct V3 { double x, y, z; };

RP_HD V3 v3(double x, double y, double z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
// Vector.tpp:11-23
RP_HD V3 add(const V3 &a, const V3 &b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
RP_HD V3 sub(const V3 &a, const V3 &b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
// Vector * scalar (Matrix.tpp:164-175); the 0 / 1 shortcuts only change signs of zeros, which never reach a
// comparison, so the plain product is used.
RP_HD V3 scal(const V3 &a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
// Vector.tpp:53-60: ((0 + x0*y0) + x1*y1) + x2*y2
RP_HD double dot(const V3 &a, const V3 &b) {
    double p = 0;
    p += a.x * b.x;
    p += a.y * b.y;
    p += a.z * b.z;
    return p;
}
// Vector.tpp:105-112
RP_HD V3 cross(const V3 &a, const V3 &b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Row-major 3x3.
struct M3 { double m[9]; };

// Matrix.tpp:105-113: r_i = ((0 + m_i0*v0) + m_i1*v1) + m_i2*v2
RP_HD V3 mulv(const double *m, const V3 &v) {
    double r0 = 0, r1 = 0, r2 = 0;
    r0 += m[0] * v.x; r0 += m[1] * v.y; r0 += m[2] * v.z;
    r1 += m[3] * v.x; r1 += m[4] * v.y; r1 += m[5] * v.z;
    r2 += m[6] * v.x; r2 += m[7] * v.y; r2 += m[8] * v.z;
    return v3(r0, r1, r2);
}

RP_HD void mulm(const double *a, const double *b, double *r) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            s += a[3 * i + 0] * b[0 + j];
            s += a[3 * i + 1] * b[3 + j];
            s += a[3 * i + 2] * b[6 + j];
            r[3 * i + j] = s;
        }
}

// Packing::fixRotationMatrix (Packing.cpp:1211-1216): three iterations of R <- 1.5 R - 0.5 R R^T R, evaluated as
// (1.5 R) - (((0.5 R) R^T) R) with Matrix.tpp's elementwise scalar products and row-by-column sums.
RP_HD void fix_rotation(double *R) {
    for (int it = 0; it < 3; it++) {
        double A[9], B[9], Rt[9], C[9], D[9];
        for (int i = 0; i < 9; i++) {
            A[i] = R[i] * 1.5;
            B[i] = R[i] * 0.5;
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) Rt[3 * i + j] = R[3 * j + i];
        mulm(B, Rt, C);
        mulm(C, R, D);
        for (int i = 0; i < 9; i++) R[i] = A[i] - D[i];
    }
}

// Matrix.tpp:352-359
RP_HD double det3(const double *a) {
    return a[0] * a[4] * a[8] + a[1] * a[5] * a[6] + a[2] * a[3] * a[7]
           - a[2] * a[4] * a[6] - a[1] * a[3] * a[8] - a[0] * a[5] * a[7];
}

// 2x2 minor with rows/columns kept in order (Matrix.tpp:296-331), det2 = b0*b3 - b1*b2
RP_HD double minor3(const double *a, int row, int col) {
    int r0 = row == 0 ? 1 : 0, r1 = row == 2 ? 1 : 2;
    int c0 = col == 0 ? 1 : 0, c1 = col == 2 ? 1 : 2;
    return a[3 * r0 + c0] * a[3 * r1 + c1] - a[3 * r0 + c1] * a[3 * r1 + c0];
}

// Matrix.tpp:258-271
RP_HD void inverse3(const double *a, double *out) {
    double det = det3(a);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            out[3 * i + j] = ((i + j) % 2 == 0) ? (minor3(a, j, i) / det) : (-minor3(a, j, i) / det);
}

// PeriodicBoundaryConditions.cpp:11-22 followed by Shape.cpp:9-12: pos += t; pos += correction(pos)
RP_HD V3 translate_wrapped(const double *dims, const double *inv, const V3 &pos, const V3 &t) {
    V3 p = add(pos, t);
    V3 rel = mulv(inv, p);
    double c[3] = {0, 0, 0};
    double r[3] = {rel.x, rel.y, rel.z};
    // The trial translation is at most half the smallest box height, so one step of each loop suffices; the
    // bound keeps a corrupted input from spinning a wavefront.
    for (int i = 0; i < 3; i++) {
        for (int it = 0; it < 4 && r[i] + c[i] < 0; it++) c[i] += 1;
        for (int it = 0; it < 4 && r[i] + c[i] >= 1; it++) c[i] -= 1;
    }
    return add(p, mulv(dims, v3(c[0], c[1], c[2])));
}

// PolysphereTraits::HardInteraction::overlapBetween (PolysphereTraits.cpp:97-106) with HardcodedTranslation
// (Packing.cpp:36-42): r = r1 + r2, then (c2 + T - c1)^2 < r * r.
RP_HD bool bead_overlap(double r1, double r2, const V3 &c1, const V3 &c2, const V3 &T) {
    const double r = r1 + r2;
    const V3 d = sub(add(c2, T), c1);
    return dot(d, d) < r * r;
}

// SegmentDistanceCalculator.h:22-85, with u = s12 - s11 and a = u.u passed in (both depend only on the first
// segment, so the sweep computes them once per trial move).
// The method and its steps are Dan Sunday's closest distance between two segments (geomalgorithms.com), which the
// reference routine is based on.  Its notice, kept with it as it asks:
//   Copyright 2001 softSurfer, 2012 Dan Sunday
//   This code may be freely used, distributed and modified for any purpose
//   providing that this copyright notice is included with it.
//   SoftSurfer makes no warranty for this code, and cannot be held
//   liable for any real or imagined damage resulting from its use.
//   Users of this code must verify correctness for their application.
RP_HD double segment_distance2(const V3 &s11, const V3 &u, double a, const V3 &s21, const V3 &s22) {
    const double EPS = 0.0000000001;
    V3 v = sub(s22, s21);
    V3 w = sub(s11, s21);
    double b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
    double D = a * c - b * b;
    double sc, sN, sD = D;
    double tc, tN, tD = D;
    if (D < EPS) {
        sN = 0.0; sD = 1.0; tN = e; tD = c;
    } else {
        sN = (b * e - c * d);
        tN = (a * e - b * d);
        if (sN < 0.0) { sN = 0.0; tN = e; tD = c; }
        else if (sN > sD) { sN = sD; tN = e + b; tD = c; }
    }
    if (tN < 0.0) {
        tN = 0.0;
        if (-d < 0.0) sN = 0.0;
        else if (-d > a) sN = sD;
        else { sN = -d; sD = a; }
    } else if (tN > tD) {
        tN = tD;
        if ((-d + b) < 0.0) sN = 0;
        else if ((-d + b) > a) sN = sD;
        else { sN = (-d + b); sD = a; }
    }
    sc = (fabs(sN) < EPS ? 0.0 : sN / sD);
    tc = (fabs(tN) < EPS ? 0.0 : tN / tD);
    V3 dP = sub(add(w, scal(u, sc)), scal(v, tc));
    return dot(dP, dP);
}

// Constants of one pair-overlap test, computed on the host exactly as the reference computes them
// (SphereTraits.cpp:64, SpherocylinderTraits.cpp:34-43).
struct PairConsts {
    int kind;            // 0 sphere, 1 spherocylinder, 2 polysphere
    int soft;            // 0 hard core; 1 WCA on the centre distance; 2 WCA on the segment distance; 3 LJ; 4 SIC
    int hardCore;        // with a soft kind: 1 = the hard test applies as well (CompoundInteraction)
    double contact2;     // sphere: pow(2r, 2); spherocylinder: 4*r*r
    double range2;       // spherocylinder: pow(2r + L, 2)
    double halfLen;      // spherocylinder: 0.5*L (cap offset, SpherocylinderTraits.cpp:24-26)
    double eps, sigma2, cut2;  // WCA: epsilon, sigma*sigma, (sigma*pow(2, 1/6))^2 (RepulsiveLennardJonesInteraction)
    double softRange2;   // centre distance^2 at and beyond which the soft energy is 0
    int nc;              // polysphere: beads per particle
    const double *beads; // polysphere: device array, per bead body-frame x, y, z and radius
    double bound;        // polysphere: max over beads of |centre| + radius, rounded up (bounding-sphere radius)
};

// Spherocylinder first-particle data that does not depend on the candidate.
struct SCFirst {
    V3 pos, capA, u;     // capA = cap centre (-1), u = capB - capA
    double a;            // u.u
};

// SpherocylinderTraits.cpp:24-26: pos + axis * ((0.5 * (+-1)) * L)
RP_HD SCFirst sc_first(const PairConsts &k, const V3 &pos, const V3 &axis) {
    SCFirst f;
    f.pos = pos;
    f.capA = add(pos, scal(axis, -k.halfLen));
    V3 capB = add(pos, scal(axis, k.halfLen));
    f.u = sub(capB, f.capA);
    f.a = dot(f.u, f.u);
    return f;
}

// SpherocylinderTraits.cpp:28-46 with pos2bc = pos2 + T (Packing.cpp:21-44)
RP_HD bool sc_overlap(const PairConsts &k, const SCFirst &f, const V3 &pos2, const V3 &axis2, const V3 &T) {
    V3 p2 = add(pos2, T);
    V3 d = sub(p2, f.pos);
    double d2 = dot(d, d);
    if (d2 < k.contact2) return true;
    if (d2 >= k.range2) return false;
    V3 c2a = add(p2, scal(axis2, -k.halfLen));
    V3 c2b = add(p2, scal(axis2, k.halfLen));
    return segment_distance2(f.capA, f.u, f.a, c2a, c2b) < k.contact2;
}

// SphereTraits.cpp:56-65 with HardcodedTranslation::getDistance2 (Packing.cpp:36-42): (p2 + T - p1).norm2()
RP_HD bool sphere_overlap(const PairConsts &k, const V3 &p1, const V3 &p2, const V3 &T) {
    V3 d = sub(add(p2, T), p1);
    return dot(d, d) < k.contact2;
}

// PeriodicBoundaryConditions::getTranslation (PeriodicBoundaryConditions.cpp:24-36): the box translation that brings
// p2 within half a box of p1 in relative coordinates (the nearest image), box * t with t integral per axis.
RP_HD V3 pbc_translation(const double *dims, const double *inv, const V3 &p1, const V3 &p2) {
    const V3 r1 = mulv(inv, p1), r2 = mulv(inv, p2);
    const double a[3] = {r1.x, r1.y, r1.z}, b[3] = {r2.x, r2.y, r2.z};
    double t[3] = {0, 0, 0};
    for (int i = 0; i < 3; i++) {
        while (b[i] + t[i] - a[i] > 0.5) t[i] -= 1;
        while (b[i] + t[i] - a[i] < -0.5) t[i] += 1;
    }
    return mulv(dims, v3(t[0], t[1], t[2]));
}

// PeriodicBoundaryConditions::getDistance2 (PeriodicBoundaryConditions.cpp:38-50): per axis |r2 - r1| in relative
// coordinates folded to at most 0.5 (the fold leaves it negative), back to absolute coordinates, squared norm.
RP_HD double pbc_distance2(const double *dims, const double *inv, const V3 &p1, const V3 &p2) {
    const V3 r1 = mulv(inv, p1), r2 = mulv(inv, p2);
    const double a[3] = {r1.x, r1.y, r1.z}, b[3] = {r2.x, r2.y, r2.z};
    double d[3];
    for (int i = 0; i < 3; i++) {
        double c = fabs(b[i] - a[i]);
        while (c > 0.5) c -= 1;
        d[i] = c;
    }
    const V3 v = mulv(dims, v3(d[0], d[1], d[2]));
    return dot(v, v);
}

// RepulsiveLennardJonesInteraction.cpp:8-16
RP_HD double wca_energy(double eps, double sigma2, double cut2, double d2) {
    if (d2 >= cut2) return 0;
    double x2 = sigma2 / d2;
    double x6 = x2 * x2 * x2;
    double x12 = x6 * x6;
    return 4 * eps * (x12 - x6) + eps;
}

// LennardJonesInteraction.cpp:8-13, zero at and beyond cut2 (infinite cut2: the reference's formula for any d)
RP_HD double lj_energy(double eps, double sigma2, double cut2, double d2) {
    if (d2 >= cut2) return 0;
    double x2 = sigma2 / d2;
    double x6 = x2 * x2 * x2;
    double x12 = x6 * x6;
    return 4 * eps * (x12 - x6);
}

// SquareInverseCoreInteraction.cpp:13-15 (sigma2 = pow(sigma, 2))
RP_HD double sic_energy(double eps, double sigma2, double d2) {
    const double v = sigma2 / d2 - 1;
    return eps * (v > 0.0 ? v : 0.0);
}

// Soft pair energy of the first particle f against pos2 + T (HardcodedTranslation).  soft 1: WCA of the centre
// distance (CentralInteraction.h:58-67); 3: Lennard-Jones and 4: square-inverse-core of the centre distance; soft 2:
// WCA of the squared distance of the two segments, the caps built as in the overlap test.  Beyond softRange2 the segments cannot be within the cut-off (segment distance >= centre
// distance - L), so the energy is 0 without the segment evaluation.
RP_HD double soft_energy(const PairConsts &k, const SCFirst &f, const V3 &pos2, const V3 &axis2, const V3 &T) {
    V3 p2 = add(pos2, T);
    V3 d = sub(p2, f.pos);
    double d2 = dot(d, d);
    if (k.soft == 1) return wca_energy(k.eps, k.sigma2, k.cut2, d2);
    if (k.soft == 3) return lj_energy(k.eps, k.sigma2, k.cut2, d2);
    if (k.soft == 4) return sic_energy(k.eps, k.sigma2, d2);
    if (d2 >= k.softRange2) return 0;
    V3 c2a = add(p2, scal(axis2, -k.halfLen));
    V3 c2b = add(p2, scal(axis2, k.halfLen));
    return wca_energy(k.eps, k.sigma2, k.cut2, segment_distance2(f.capA, f.u, f.a, c2a, c2b));
}

// Philox4x32-10 (Salmon et al., SC'11), counter-based: one call maps (counter, key) to four 32-bit words.
struct Philox {
    uint32_t c0, c1, c2, c3, k0, k1;
    uint32_t b0, b1, b2, b3;  // current output block (separate members: no dynamically indexed array)
    int used;
};

RP_HD void philox_round(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    uint64_t p0 = (uint64_t)M0 * c0;
    uint64_t p1 = (uint64_t)M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
}

RP_HD void philox_block(const uint32_t ctr[4], uint32_t key0, uint32_t key1, uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key0, k1 = key1;
    for (int r = 0; r < 10; r++) {
        philox_round(c0, c1, c2, c3, k0, k1);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

RP_HD Philox philox_init(uint64_t seed, uint32_t s0, uint32_t s1, uint32_t s2) {
    Philox p;
    p.k0 = (uint32_t)seed; p.k1 = (uint32_t)(seed >> 32);
    p.c0 = 0; p.c1 = s0; p.c2 = s1; p.c3 = s2;
    p.used = 4;
    return p;
}

RP_HD uint32_t philox_u32(Philox &p) {
    if (p.used == 4) {
        uint32_t ctr[4] = {p.c0, p.c1, p.c2, p.c3}, out[4];
        philox_block(ctr, p.k0, p.k1, out);
        p.b0 = out[0]; p.b1 = out[1]; p.b2 = out[2]; p.b3 = out[3];
        p.c0++;
        p.used = 0;
    }
    uint32_t v = p.used == 0 ? p.b0 : (p.used == 1 ? p.b1 : (p.used == 2 ? p.b2 : p.b3));
    p.used++;
    return v;
}

// Uniform double in [0, 1) with 53 random bits.
RP_HD double philox_unit(Philox &p) {
    uint64_t hi = philox_u32(p), lo = philox_u32(p);
    uint64_t bits = ((hi << 32) | lo) >> 11;
    return (double)bits * (1.0 / 9007199254740992.0);
}

// Uniform in [a, b) as std::uniform_real_distribution maps a unit draw: a + u*(b - a).
RP_HD double philox_uniform(Philox &p, double a, double b) { return a + philox_unit(p) * (b - a); }

// Uniform integer in [0, n) (multiply-shift; bias < n / 2^32).
RP_HD uint32_t philox_below(Philox &p, uint32_t n) { return (uint32_t)(((uint64_t)philox_u32(p) * n) >> 32); }

// ---------------------------------------------------------------------------------------------------------------
// Guarded FP32 pair decisions.  Distances are evaluated in FP32 on coordinates taken relative to a reference point,
// each with a bound on its error against the reference's FP64 value; a pair whose FP32 value lies farther from the
// decision threshold than the bound is decided in FP32, any other pair is left to the FP64 path, so the decision is
// always the FP64 one.  Every bound is per component and uses u = 2^-24 (FP32 unit roundoff) with a 1 % margin.
//   Centre distance: d2f = |fj - fm|^2 from FP32 coordinates fm, fj that are roundings of FP64 differences; the bound
//   covers both input roundings, the subtraction and the three-term sum.
//   Segment distance (spherocylinders): the squared distance between points s, t of the two axis segments is a convex
//   quadratic on [0, 1]^2, so at ANY (s, t) its value bounds the minimum from above and its value plus the minimum of
//   its linear part over the corners bounds the minimum from below.  (s, t) come from the FP32 form of the reference's
//   algorithm (SegmentDistanceCalculator.h:22-85) but need not be exact; only the evaluation at (s, t) and the
//   perturbation of the four cap points by the FP32 roundings (which moves the minimum distance by at most twice the
//   largest cap displacement) enter the bounds.  Nearly parallel pairs (D < 1e-3 a c) go to FP64: there the reference's
//   own parallel branch (D < 1e-10) need not return the minimum.
constexpr float GUARD_U = 5.9604645e-08f * 1.01f;  // 2^-24 with margin
constexpr double GUARD_SLACK64 = 1e-6;             // the reference's FP64 rounding and EPS-clamped parameters

enum GuardResult : int { GUARD_NO = 0, GUARD_OVERLAP = 1, GUARD_UNDECIDED = 2 };

struct F3 { float x, y, z; };
RP_HD F3 f3(float x, float y, float z) { F3 r; r.x = x; r.y = y; r.z = z; return r; }
RP_HD F3 fadd(const F3 &a, const F3 &b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
RP_HD F3 fsub(const F3 &a, const F3 &b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
RP_HD F3 fscal(const F3 &a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
RP_HD float fdot(const F3 &a, const F3 &b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
RP_HD F3 fabs3(const F3 &a) { return f3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
// FP32 rounding of an FP64 coordinate and the bound on |result - exact|: u |x| (the FP64 input is exact data, or an
// FP64 difference whose own rounding, 2^-53 |x|, the margin covers)
RP_HD F3 to_f3(const V3 &a) { return f3((float)a.x, (float)a.y, (float)a.z); }
RP_HD F3 round_err(const F3 &a) { return fscal(fabs3(a), GUARD_U); }

// Centre-distance stage.  fm, fj: FP32 roundings of (pos - r0); em, ej their error bounds.  Returns d2f and sets
// `bound` so that |d2f - d2_64| <= bound for the reference's FP64 d2 of the same pair.
RP_HD float guard_centre(const F3 &fm, const F3 &em, const F3 &fj, const F3 &ej, float &bound) {
    const F3 d = fsub(fj, fm);
    const F3 ad = fabs3(d);
    const F3 e = fadd(fadd(em, ej), fscal(ad, GUARD_U));  // inputs + the subtraction
    const float d2 = fdot(d, d);
    bound = ((2.f * ad.x + e.x) * e.x + (2.f * ad.y + e.y) * e.y + (2.f * ad.z + e.z) * e.z) * 1.01f +
            3.1f * GUARD_U * d2 + 1e-12f;
    return d2;
}

// Segment stage for a spherocylinder pair whose centre distance is certainly within (contact, range).  fm, am: FP32
// centre (relative) and axis of the first particle, em its position error bound; fj, aj, ej likewise; h = L / 2 in
// FP32 with its rounding error eh.  Returns GUARD_NO, GUARD_OVERLAP or GUARD_UNDECIDED; d2s receives the FP32
// squared segment distance at the chosen (s, t) (for error statistics).
RP_HD int guard_segment(double contact2, const F3 &fm, const F3 &am, const F3 &em, const F3 &fj, const F3 &aj,
                        const F3 &ej, float h, float eh, float &d2s) {
    // cap points in FP32 and the bound on their displacement from the reference's caps (relative to r0)
    const F3 ha = fscal(am, h), hj = fscal(aj, h);
    const F3 ca = fsub(fm, ha), cb = fadd(fm, ha), da = fsub(fj, hj), db = fadd(fj, hj);
    auto capErr = [&](const F3 &c, const F3 &p, const F3 &ep, const F3 &a, const F3 &hh) {
        // position, axis rounding times h, h rounding times |axis|, the product and the sum
        const F3 aa = fabs3(a);
        const float x = ep.x + GUARD_U * aa.x * h + eh * aa.x + GUARD_U * fabsf(hh.x) + GUARD_U * fabsf(c.x);
        const float y = ep.y + GUARD_U * aa.y * h + eh * aa.y + GUARD_U * fabsf(hh.y) + GUARD_U * fabsf(c.y);
        const float z = ep.z + GUARD_U * aa.z * h + eh * aa.z + GUARD_U * fabsf(hh.z) + GUARD_U * fabsf(c.z);
        (void)p;
        return (x + y + z) * 1.01f + 1e-9f;  // >= the Euclidean displacement
    };
    const float delta = fmaxf(fmaxf(capErr(ca, fm, em, am, ha), capErr(cb, fm, em, am, ha)),
                              fmaxf(capErr(da, fj, ej, aj, hj), capErr(db, fj, ej, aj, hj)));
    // the reference's parametrisation: S1 = ca + s u, S2 = da + t v, w = ca - da
    const F3 u = fsub(cb, ca), v = fsub(db, da), w = fsub(ca, da);
    const float a = fdot(u, u), b = fdot(u, v), c = fdot(v, v), d = fdot(u, w), e = fdot(v, w);
    const float D = a * c - b * b;
    if (!(D >= 1e-3f * a * c)) return GUARD_UNDECIDED;  // nearly parallel (or degenerate)
    // FP32 form of SegmentDistanceCalculator.h:22-85 (general branch); any (s, t) in [0, 1]^2 gives valid bounds
    float sN = b * e - c * d, sD = D, tN = a * e - b * d, tD = D;
    if (sN < 0.f) { sN = 0.f; tN = e; tD = c; }
    else if (sN > sD) { sN = sD; tN = e + b; tD = c; }
    if (tN < 0.f) {
        tN = 0.f;
        if (-d < 0.f) sN = 0.f;
        else if (-d > a) sN = sD;
        else { sN = -d; sD = a; }
    } else if (tN > tD) {
        tN = tD;
        if ((-d + b) < 0.f) sN = 0.f;
        else if ((-d + b) > a) sN = sD;
        else { sN = -d + b; sD = a; }
    }
    const float sc = fminf(fmaxf(sN / sD, 0.f), 1.f), tc = fminf(fmaxf(tN / tD, 0.f), 1.f);
    // residual at (sc, tc) and its error against the exact residual of the FP32 cap points
    const F3 wsu = fadd(w, fscal(u, sc));
    const F3 r = fsub(wsu, fscal(v, tc));
    const F3 aw = fabs3(w), au = fabs3(u), av = fabs3(v), ar = fabs3(r);
    const F3 er = fscal(fadd(fadd(fadd(aw, au), fadd(av, fabs3(fscal(u, sc)))),
                             fadd(fadd(fabs3(wsu), fabs3(fscal(v, tc))), ar)),
                        GUARD_U * 1.01f);
    const float f = fdot(r, r);
    const float ef = ((2.f * ar.x + er.x) * er.x + (2.f * ar.y + er.y) * er.y + (2.f * ar.z + er.z) * er.z) * 1.01f +
                     3.1f * GUARD_U * f;
    d2s = f;
    // gradient of |w + s u - t v|^2 at (sc, tc): (2 u.r, -2 v.r), and the minimum of its linear part over [0, 1]^2
    const float gs = 2.f * fdot(u, r), gt = -2.f * fdot(v, r);
    const float egs = 2.f * (fdot(au, er) + fdot(fadd(ar, er), fscal(au, GUARD_U)) + 3.1f * GUARD_U * fdot(au, ar));
    const float egt = 2.f * (fdot(av, er) + fdot(fadd(ar, er), fscal(av, GUARD_U)) + 3.1f * GUARD_U * fdot(av, ar));
    const float ls = fminf(-gs * sc, gs * (1.f - sc)), lt = fminf(-gt * tc, gt * (1.f - tc));
    // bounds on the minimum squared distance of the FP32 cap geometry, in FP64 from here on
    const double ub = (double)f + (double)ef;
    const double lb = (double)f - (double)ef + (double)ls + (double)lt - (double)egs - (double)egt -
                      2.1 * (double)GUARD_U * ((double)fabsf(gs) + (double)fabsf(gt)) -
                      4.0 * (double)GUARD_U * ((double)f + (double)fabsf(ls) + (double)fabsf(lt));
    // minimum distance of the reference geometry: within 2 delta of the FP32 geometry's
    const double dd = 2.0 * (double)delta;
    const double hiDist = sqrt(ub) + dd;
    if (hiDist * hiDist + GUARD_SLACK64 < contact2) return GUARD_OVERLAP;
    if (lb > 0) {
        const double loDist = sqrt(lb) - dd;
        if (loDist > 0 && loDist * loDist - GUARD_SLACK64 >= contact2) return GUARD_NO;
    }
    return GUARD_UNDECIDED;
}

}  // namespace rp
