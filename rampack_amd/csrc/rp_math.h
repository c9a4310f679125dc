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

// SegmentDistanceCalculator.h:22-85, with u = s12 - s11 and a = u.u passed in (both depend only on the first
// segment, so the sweep computes them once per trial move).
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
    int kind;            // 0 sphere, 1 spherocylinder
    int soft;            // 0 hard core; 1 WCA on the centre distance; 2 WCA on the segment distance
    double contact2;     // sphere: pow(2r, 2); spherocylinder: 4*r*r
    double range2;       // spherocylinder: pow(2r + L, 2)
    double halfLen;      // spherocylinder: 0.5*L (cap offset, SpherocylinderTraits.cpp:24-26)
    double eps, sigma2, cut2;  // WCA: epsilon, sigma*sigma, (sigma*pow(2, 1/6))^2 (RepulsiveLennardJonesInteraction)
    double softRange2;   // centre distance^2 at and beyond which the soft energy is 0
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

// RepulsiveLennardJonesInteraction.cpp:8-16
RP_HD double wca_energy(double eps, double sigma2, double cut2, double d2) {
    if (d2 >= cut2) return 0;
    double x2 = sigma2 / d2;
    double x6 = x2 * x2 * x2;
    double x12 = x6 * x6;
    return 4 * eps * (x12 - x6) + eps;
}

// Soft pair energy of the first particle f against pos2 + T (HardcodedTranslation).  soft 1: WCA of the centre
// distance (CentralInteraction.h:58-67); soft 2: WCA of the squared distance of the two segments, the caps built as
// in the overlap test.  Beyond softRange2 the segments cannot be within the cut-off (segment distance >= centre
// distance - L), so the energy is 0 without the segment evaluation.
RP_HD double soft_energy(const PairConsts &k, const SCFirst &f, const V3 &pos2, const V3 &axis2, const V3 &T) {
    V3 p2 = add(pos2, T);
    V3 d = sub(p2, f.pos);
    double d2 = dot(d, d);
    if (k.soft == 1) return wca_energy(k.eps, k.sigma2, k.cut2, d2);
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

}  // namespace rp
