// oxr_device.h -- device-side vector math, Philox RNG, tables and CSG intersection.
//
// Float semantics are deliberate: sources are compiled with -ffp-contract=off and
// correctly rounded fp32 divide/sqrt, and log/sin/cos/atan2 are evaluated in fp64 and
// rounded, so every fp32 result is a function of its inputs alone.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define OXD __device__ __forceinline__

namespace oxd {

// ------------------------------------------------------------------ vectors
struct f3 { float x, y, z; };
OXD f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
OXD f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
OXD f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
OXD f3 neg(f3 a) { return mk(-a.x, -a.y, -a.z); }
OXD f3 mul(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
OXD f3 mulr(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
OXD f3 divs(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
OXD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
OXD f3 cross(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
OXD float len(f3 a) { return sqrtf(dot(a, a)); }
OXD f3 normalize(f3 a) { float inv = 1.0f / sqrtf(dot(a, a)); return mulr(a, inv); }

// Natural log in single precision for the photon loop: the argument reduction and
// minimax polynomial of FreeBSD msun's e_logf.c (x = 2^k (1+f), sqrt(2)/2 <= 1+f < sqrt(2),
// s = f/(2+f), log(1+f) = f - f^2/2 + s (f^2/2 + R(s^2))), in IEEE single-precision
// operations only, so the oracle's restatement gives the same bits. Faithfully rounded
// (max 0.83 ulp over the 2^24 Philox uniforms). Zero, negative, subnormal, infinite
// and NaN arguments take the double-precision routine.
OXD float xlog(float x) {
    uint32_t ix = __float_as_uint(x);
    if (ix >= 0x7f800000u || ix < 0x00800000u) return (float)log((double)x);
    const float ln2_hi = 6.9313812256e-01f, ln2_lo = 9.0580006145e-06f;
    const float Lg1 = 0.66666662693f, Lg2 = 0.40000972152f, Lg3 = 0.28498786688f, Lg4 = 0.24279078841f;
    int k = int(ix >> 23) - 127;
    uint32_t m = ix & 0x007fffffu;
    uint32_t i = (m + 0x004afb20u) & 0x00800000u;          // set when 1.m >= sqrt(2): use (1.m)/2
    float xm = __uint_as_float(m | (i ^ 0x3f800000u));
    k += int(i >> 23);
    float f = xm - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s, w = z * z;
    float t1 = w * (Lg2 + w * Lg4);
    float t2 = z * (Lg1 + w * Lg3);
    float R = t2 + t1;
    float hfsq = 0.5f * f * f;
    float dk = float(k);
    return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
}
OXD float xsin(float x) { return (float)sin((double)x); }
OXD float xcos(float x) { return (float)cos((double)x); }
OXD float xatan2(float y, float x) { return (float)atan2((double)y, (double)x); }
constexpr float PIf = 3.14159265358979323846f;

// ------------------------------------------------------------------ flags
enum : uint32_t {
    CERENKOV = 1u << 0, SCINTILLATION = 1u << 1, TORCH = 1u << 2, BULK_ABSORB = 1u << 3,
    BULK_REEMIT = 1u << 4, BULK_SCATTER = 1u << 5, SURFACE_DETECT = 1u << 6,
    SURFACE_ABSORB = 1u << 7, SURFACE_DREFLECT = 1u << 8, SURFACE_SREFLECT = 1u << 9,
    BOUNDARY_REFLECT = 1u << 10, BOUNDARY_TRANSMIT = 1u << 11
};
enum { BREAK = 1, CONTINUE = 2, BOUNDARY = 3 };
enum { CSG_ZERO = 0, CSG_NODE = 11, CSG_CONTIGUOUS = 11, CSG_DISCONTIGUOUS = 12, CSG_OVERLAP = 13, CSG_LEAF = 101, CSG_SPHERE = 101, CSG_ZSPHERE = 103,
       CSG_CYLINDER = 105, CSG_CONE = 108, CSG_BOX3 = 110, CSG_CONVEXPOLYHEDRON = 112,
       CSG_HYPERBOLOID = 117, CSG_PHICUT = 121, CSG_HALFSPACE = 125 };
constexpr float RT_DEFAULT_MAX = 1.e27f;   // CSG/csg_intersect_leaf_head.h:10

// ------------------------------------------------------------------ Philox4x32-10
// Counter-based generator of Salmon et al. (SC'11). Stream of photon p: key = seed,
// counter = (block lo, block hi, p lo, p hi), uniform j at block (start+j)/4, lane (start+j)%4.
// The four outputs of the current block are held in scalars b0..b3 and consumed by
// shifting, so the state never needs an indexed (memory-resident) array.
struct Rng {
    uint32_t k0, k1;
    uint32_t p0, p1;          // photon index (subsequence)
    uint64_t pos;             // absolute uniform index of the next uniform
    uint32_t b0, b1, b2, b3;  // unconsumed outputs of the current block, b0 next
    int have;                 // how many of b0..b3 are valid
    uint32_t c0, c1, c2, c3;  // the following block, computed ahead by rng_topup
    int have2;                // 4 when c0..c3 hold it, else 0
    const float* pre;         // precooked row or null
    int32_t npre, cur;
    uint32_t overflow;
};

// Philox4x32-10. Each round's 32x32 products come from one 64-bit multiply
// (v_mad_u64_u32 gives both halves). The key is the same for every photon of a launch:
// readfirstlane keeps its schedule in scalar registers.
OXD void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                uint32_t& o0, uint32_t& o1, uint32_t& o2, uint32_t& o3) {
    k0 = __builtin_amdgcn_readfirstlane(k0);
    k1 = __builtin_amdgcn_readfirstlane(k1);
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0, n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = uint32_t(p1); c2 = n2; c3 = uint32_t(p0);
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    o0 = c0; o1 = c1; o2 = c2; o3 = c3;
}

OXD void rng_init(Rng& r, uint64_t seed, uint64_t offset, uint64_t skip, uint64_t event, uint64_t photon) {
    r.k0 = uint32_t(seed); r.k1 = uint32_t(seed >> 32);
    r.p0 = uint32_t(photon); r.p1 = uint32_t(photon >> 32);
    r.pos = offset + skip * event;
    r.have = 0; r.have2 = 0;
    r.b0 = r.b1 = r.b2 = r.b3 = 0u;
    r.c0 = r.c1 = r.c2 = r.c3 = 0u;
    r.pre = nullptr; r.npre = 0; r.cur = 0; r.overflow = 0;
}

// continue a stream at uniform number pos (the counter-based stream needs nothing else)
OXD void rng_resume(Rng& r, uint64_t seed, uint64_t pos, uint64_t photon) {
    r.k0 = uint32_t(seed); r.k1 = uint32_t(seed >> 32);
    r.p0 = uint32_t(photon); r.p1 = uint32_t(photon >> 32);
    r.pos = pos;
    r.have = 0; r.have2 = 0;
    r.b0 = r.b1 = r.b2 = r.b3 = 0u;
    r.c0 = r.c1 = r.c2 = r.c3 = 0u;
    r.pre = nullptr; r.npre = 0; r.cur = 0; r.overflow = 0;
}

// refill an empty buffer: from the block computed ahead, else from Philox
OXD void rng_refill(Rng& r) {
    if (r.have2) {                             // the block after a buffer always starts aligned
        r.b0 = r.c0; r.b1 = r.c1; r.b2 = r.c2; r.b3 = r.c3;
        r.have = 4; r.have2 = 0;
        return;
    }
    uint64_t block = r.pos >> 2;
    philox(uint32_t(block), uint32_t(block >> 32), r.p0, r.p1, r.k0, r.k1, r.b0, r.b1, r.b2, r.b3);
    uint32_t skip = uint32_t(r.pos & 3);      // only the first block of a stream can start mid-block
    if (skip >= 1u) { r.b0 = r.b1; r.b1 = r.b2; r.b2 = r.b3; }
    if (skip >= 2u) { r.b0 = r.b1; r.b1 = r.b2; }
    if (skip >= 3u) { r.b0 = r.b1; }
    r.have = 4 - int(skip);
}

// Compute the block after the buffered uniforms now, so that the next five to eight
// uniforms need no Philox. Called where every lane of a branch is about to draw (the
// start of a bounce's physics, the diffuse-reflection rejection loops): the Philox rounds
// then run once for the branch instead of inside divergent loops whenever some lane's
// buffer empties. The values drawn are unchanged; only when they are computed moves.
OXD void rng_topup(Rng& r) {
    if (r.pre || r.have2) return;
    if (r.have == 0) rng_refill(r);
    uint64_t next = (r.pos + uint64_t(r.have)) >> 2;
    philox(uint32_t(next), uint32_t(next >> 32), r.p0, r.p1, r.k0, r.k1, r.c0, r.c1, r.c2, r.c3);
    r.have2 = 4;
}

OXD float uniform(Rng& r) {
    if (r.pre) {
        if (r.cur >= r.npre) { r.cur = 0; r.overflow++; }
        return r.pre[r.cur++];
    }
    if (r.have == 0) rng_refill(r);
    uint32_t x = r.b0;
    r.b0 = r.b1; r.b1 = r.b2; r.b2 = r.b3;
    r.have--;
    r.pos++;
    return float((x >> 8) + 1u) * (1.0f / 16777216.0f);
}

// ------------------------------------------------------------------ tables
// bnd texture emulation: linear filter, normalised coordinates, wrap (qudarap/qbnd.h:103-125)
struct Tables {
    const float4* bnd;      // rows of num_wl float4, row = 2*line + k
    const uint4* optical;   // (num_bnd*4)
    const float* icdf;      // (rows, width)
    int num_wl;
    float wl_min, wl_step;
    int icdf_width, hd_factor;
};

// the wavelength coordinate of a lookup: the two texels and their weights (shared by every
// row looked up at one wavelength)
struct TexX { int m0, m1; float a, b; };
OXD TexX tex_x(const Tables& t, float nm) {
    const int nx = t.num_wl;
    float fx = (nm - t.wl_min) / t.wl_step;
    float x = (fx + 0.5f) / float(nx);
    float xf = x * float(nx) - 0.5f;
    float fl = floorf(xf);
    TexX X;
    X.a = xf - fl;
    X.b = 1.0f - X.a;
    int i0 = int(fl);
    X.m0 = i0 % nx; if (X.m0 < 0) X.m0 += nx;
    X.m1 = X.m0 + 1; if (X.m1 == nx) X.m1 = 0;
    return X;
}
OXD float4 boundary_fetch(const Tables& t, const TexX& X, unsigned line, unsigned k) {
    const float4* row = t.bnd + size_t(2u * line + k) * t.num_wl;
    float4 v0 = row[X.m0], v1 = row[X.m1];
    const float a = X.a, b = X.b;
    return make_float4(b * v0.x + a * v1.x, b * v0.y + a * v1.y, b * v0.z + a * v1.z, b * v0.w + a * v1.w);
}
OXD float4 boundary_lookup(const Tables& t, float nm, unsigned line, unsigned k) {
    return boundary_fetch(t, tex_x(t, nm), line, k);
}

OXD float icdf_lookup(const Tables& t, float x, int row) {
    const int nx = t.icdf_width;
    float xf = x * float(nx) - 0.5f;
    float fl = floorf(xf);
    float a = xf - fl;
    int i0 = int(fl);
    int m0 = i0 % nx; if (m0 < 0) m0 += nx;
    int m1 = m0 + 1; if (m1 == nx) m1 = 0;
    const float* r = t.icdf + size_t(row) * nx;
    return (1.0f - a) * r[m0] + a * r[m1];
}

// qscint::wavelength_hd0/10/20 (qudarap/qscint.h:150-228)
OXD float scint_wavelength(const Tables& t, float u0) {
    if (!t.icdf) return 0.f;
    if (t.hd_factor == 20) {
        if (u0 < 0.05f) return icdf_lookup(t, u0 * 20.f, 1);
        if (u0 > 0.95f) return icdf_lookup(t, (u0 - 0.95f) * 20.f, 2);
        return icdf_lookup(t, u0, 0);
    }
    if (t.hd_factor == 10) {
        if (u0 < 0.1f) return icdf_lookup(t, u0 * 10.f, 1);
        if (u0 > 0.9f) return icdf_lookup(t, (u0 - 0.9f) * 10.f, 2);
        return icdf_lookup(t, u0, 0);
    }
    return icdf_lookup(t, u0, 0);
}

// ------------------------------------------------------------------ CSG
struct Is { float x, y, z, w; };

// CSGNode as four float4: q0 params, q1 (z1, z2, boundary, index), q2/q3 aabb, q3.z typecode,
// q3.w gtransformIdx|complement (CSG/CSGNode.h:67-98)
struct NodeV { float4 q0, q1, q3; };
OXD NodeV load_node(const float4* node, int i) {
    NodeV n;
    n.q0 = node[4 * i + 0];
    n.q1 = node[4 * i + 1];
    n.q3 = node[4 * i + 3];
    return n;
}
OXD unsigned typecode(const NodeV& n) { return __float_as_uint(n.q3.z); }

// csg_robust_quadratic_roots.h:95-134
OXD void rq_roots(float& t1, float& t2, float& sdisc, float d, float b, float c) {
    float disc = b * b - d * c;
    sdisc = disc > 0.f ? sqrtf(disc) : 0.f;
    float q = b > 0.f ? -(b + sdisc) : -(b - sdisc);
    float r1 = q / d, r2 = c / q;
    t1 = fminf(r1, r2); t2 = fmaxf(r1, r2);
}
OXD void rq_roots64(double& t1, double& t2, double& sdisc, double d, double b, double c) {
    double disc = b * b - d * c;
    sdisc = disc > 0.0 ? sqrt(disc) : 0.0;
    double q = b > 0.0 ? -(b + sdisc) : -(b - sdisc);
    double r1 = q / d, r2 = c / q;
    t1 = fmin(r1, r2); t2 = fmax(r1, r2);
}
OXD void rq_roots_disq(float t_min, float& t1, float& t2, float& sdisc, float d, float b, float c) {
    float disc = b * b - d * c;
    sdisc = disc > 0.f ? sqrtf(disc) : 0.f;
    float q = b > 0.f ? -(b + sdisc) : -(b - sdisc);
    float r1 = sdisc > 0.f ? q / d : t_min;
    float r2 = sdisc > 0.f ? c / q : t_min;
    t1 = fminf(r1, r2); t2 = fmaxf(r1, r2);
}
OXD void rq_roots_disq64(double t_min, double& t1, double& t2, double& sdisc, double d, double b, double c) {
    double disc = b * b - d * c;
    sdisc = disc > 0.0 ? sqrt(disc) : 0.0;
    double q = b > 0.0 ? -(b + sdisc) : -(b - sdisc);
    double r1 = sdisc > 0.0 ? q / d : t_min;
    double r2 = sdisc > 0.0 ? c / q : t_min;
    t1 = fmin(r1, r2); t2 = fmax(r1, r2);
}

// csg_intersect_leaf_sphere.h:15-55
template <bool FP64>
OXD bool leaf_sphere(Is& is, const NodeV& n, float t_min, f3 o, f3 d) {
    f3 O = sub(o, mk(n.q0.x, n.q0.y, n.q0.z));
    float radius = n.q0.w;
    float root1, root2, sdisc;
    if (FP64) {
        double ox = O.x, oy = O.y, oz = O.z, dx = d.x, dy = d.y, dz = d.z;
        double b = ox * dx + oy * dy + oz * dz;
        double c = (ox * ox + oy * oy + oz * oz) - double(radius) * double(radius);
        double dd = dx * dx + dy * dy + dz * dz;
        double r1, r2, sd;
        rq_roots64(r1, r2, sd, dd, b, c);
        root1 = float(r1); root2 = float(r2); sdisc = float(sd);
    } else {
        float b = dot(O, d);
        float c = dot(O, O) - radius * radius;
        float dd = dot(d, d);
        rq_roots(root1, root2, sdisc, dd, b, c);
    }
    float t = sdisc > 0.f ? (root1 > t_min ? root1 : root2) : t_min;
    bool valid = t > t_min;
    if (valid) {
        is.x = (O.x + t * d.x) / radius;
        is.y = (O.y + t * d.y) / radius;
        is.z = (O.z + t * d.z) / radius;
        is.w = t;
    }
    return valid;
}

// csg_intersect_leaf_hyperboloid.h:18-106: one-sheet hyperboloid x^2+y^2 = r0^2((z/zf)^2+1)
// between planes z1 < z2 (q0 = r0, zf, z1, z2); float roots in both precision modes
OXD bool leaf_hyperboloid(Is& is, const NodeV& n, float t_min, f3 o, f3 d) {
    const float zero = 0.f, one = 1.f;
    const float r0 = n.q0.x, zf = n.q0.y, z1 = n.q0.z, z2 = n.q0.w;
    const float rr0 = r0 * r0;
    const float z1s = z1 / zf, z2s = z2 / zf;
    const float rr1 = rr0 * (z1s * z1s + one), rr2 = rr0 * (z2s * z2s + one);
    const float A = -rr0 / (zf * zf);
    const float B = -rr0;
    const float sx = d.x, sy = d.y, sz = d.z, ox = o.x, oy = o.y, oz = o.z;
    const float dd = sx * sx + sy * sy + A * sz * sz;
    const float b = ox * sx + oy * sy + A * oz * sz;
    const float c = ox * ox + oy * oy + A * oz * oz + B;
    const float disc = b * b - dd * c;
    float t1hyp, t2hyp, sdisc;
    rq_roots(t1hyp, t2hyp, sdisc, dd, b, c);
    const float h1z = oz + t1hyp * sz, h2z = oz + t2hyp * sz;
    const float osz = one / sz;
    const float t2cap = (z2 - oz) * osz, t1cap = (z1 - oz) * osz;
    const f3 c1 = add(o, mul(t1cap, d)), c2 = add(o, mul(t2cap, d));
    const float crr1 = c1.x * c1.x + c1.y * c1.y, crr2 = c2.x * c2.x + c2.y * c2.y;
    const float tc0 = t1hyp > t_min && disc > zero && h1z > z1 && h1z < z2 ? t1hyp : RT_DEFAULT_MAX;
    const float tc1 = t2hyp > t_min && disc > zero && h2z > z1 && h2z < z2 ? t2hyp : RT_DEFAULT_MAX;
    const float tc2 = t2cap > t_min && crr2 < rr2 ? t2cap : RT_DEFAULT_MAX;
    const float tc3 = t1cap > t_min && crr1 < rr1 ? t1cap : RT_DEFAULT_MAX;
    const float t_cand = fminf(fminf(tc0, tc1), fminf(tc2, tc3));
    const bool valid = t_cand > t_min && t_cand < RT_DEFAULT_MAX;
    if (valid) {
        is.w = t_cand;
        if (t_cand == t1hyp || t_cand == t2hyp) {
            const f3 p = add(o, mul(t_cand, d));
            const f3 nn = normalize(mk(p.x, p.y, A * p.z));
            is.x = nn.x; is.y = nn.y; is.z = nn.z;
        } else {
            is.x = zero; is.y = zero; is.z = t_cand == t1cap ? -one : one;
        }
    }
    return valid;
}

// csg_intersect_leaf_zsphere.h:45-151
template <bool FP64>
OXD bool leaf_zsphere(Is& is, const NodeV& n, float t_min, f3 ro, f3 rd) {
    f3 center = mk(n.q0.x, n.q0.y, n.q0.z);
    f3 O = sub(ro, center);
    f3 D = rd;
    float radius = n.q0.w;
    float b = dot(O, D);
    float c = dot(O, O) - radius * radius;
    if (c > 0.f && b > 0.f) return false;
    float zmax = center.z + n.q1.y;
    float zmin = center.z + n.q1.x;
    float d = dot(D, D);
    float t1sph, t2sph, sdisc;
    if (FP64) {
        double ox = O.x, oy = O.y, oz = O.z, dx = D.x, dy = D.y, dz = D.z;
        double bd = ox * dx + oy * dy + oz * dz;
        double cd = (ox * ox + oy * oy + oz * oz) - double(radius) * double(radius);
        double dd = dx * dx + dy * dy + dz * dz;
        double r1, r2, sd;
        rq_roots64(r1, r2, sd, dd, bd, cd);
        t1sph = float(r1); t2sph = float(r2); sdisc = float(sd);
    } else {
        rq_roots(t1sph, t2sph, sdisc, d, b, c);
    }
    float z1sph = ro.z + t1sph * rd.z;
    float z2sph = ro.z + t2sph * rd.z;
    float idz = 1.f / rd.z;
    float t_QCAP = (zmax - ro.z) * idz;
    float t_PCAP = (zmin - ro.z) * idz;
    float t1cap = fminf(t_QCAP, t_PCAP);
    float t2cap = fmaxf(t_QCAP, t_PCAP);
    if (t1cap < t1sph || t1cap > t2sph) t1cap = t_min;
    if (t2cap < t1sph || t2cap > t2sph) t2cap = t_min;
    float t = t_min;
    if (sdisc > 0.f) {
        if (t1sph > t_min && z1sph > zmin && z1sph <= zmax) t = t1sph;
        else if (t1cap > t_min) t = t1cap;
        else if (t2cap > t_min) t = t2cap;
        else if (t2sph > t_min && z2sph > zmin && z2sph <= zmax) t = t2sph;
    }
    bool valid = t > t_min;
    if (valid) {
        is.w = t;
        if (t == t1sph || t == t2sph) {
            is.x = (O.x + t * D.x) / radius;
            is.y = (O.y + t * D.y) / radius;
            is.z = (O.z + t * D.z) / radius;
        } else {
            is.x = 0.f; is.y = 0.f;
            is.z = t == t_PCAP ? -1.f : 1.f;
        }
    }
    return valid;
}

// csg_intersect_leaf_cylinder.h:34-82
template <bool FP64>
OXD bool leaf_cylinder(Is& is, const NodeV& n, float t_min, f3 o, f3 v) {
    float r = n.q0.w, z1 = n.q1.x, z2 = n.q1.y;
    float ox = o.x, oy = o.y, oz = o.z, vx = v.x, vy = v.y, vz = v.z;
    float r2 = r * r;
    float t_near, t_far, sdisc;
    if (FP64) {
        double ad = double(vx) * vx + double(vy) * vy, bd = double(ox) * vx + double(oy) * vy;
        double cd = double(ox) * ox + double(oy) * oy - double(r) * double(r);
        double tn, tf, sd;
        rq_roots_disq64(double(t_min), tn, tf, sd, ad, bd, cd);
        t_near = float(tn); t_far = float(tf); sdisc = float(sd);
    } else {
        float a = vx * vx + vy * vy;
        float b = ox * vx + oy * vy;
        float c = ox * ox + oy * oy - r2;
        rq_roots_disq(t_min, t_near, t_far, sdisc, a, b, c);
    }
    float z_near = oz + t_near * vz;
    float z_far = oz + t_far * vz;
    float t_z1cap = (z1 - oz) / vz;
    float r2_z1cap = (ox + t_z1cap * vx) * (ox + t_z1cap * vx) + (oy + t_z1cap * vy) * (oy + t_z1cap * vy);
    float t_z2cap = (z2 - oz) / vz;
    float r2_z2cap = (ox + t_z2cap * vx) * (ox + t_z2cap * vx) + (oy + t_z2cap * vy) * (oy + t_z2cap * vy);
    const float INF = __int_as_float(0x7f800000);
    float t = INF;
    if (t_near > t_min && z_near > z1 && z_near < z2 && t_near < t) t = t_near;
    if (t_far > t_min && z_far > z1 && z_far < z2 && t_far < t) t = t_far;
    if (t_z1cap > t_min && r2_z1cap <= r2 && t_z1cap < t) t = t_z1cap;
    if (t_z2cap > t_min && r2_z2cap <= r2 && t_z2cap < t) t = t_z2cap;
    bool valid = t > t_min && t < INF;
    if (valid) {
        bool sheet = (t == t_near || t == t_far);
        is.x = sheet ? (ox + t * vx) / r : 0.f;
        is.y = sheet ? (oy + t * vy) / r : 0.f;
        is.z = sheet ? 0.f : (t == t_z1cap ? -1.f : 1.f);
        is.w = t;
    }
    return valid;
}

// csg_intersect_leaf_box3.h:64-157
OXD bool leaf_box3(Is& is, const NodeV& n, float t_min, f3 o, f3 d) {
    f3 bmin = mk(-n.q0.x / 2.f, -n.q0.y / 2.f, -n.q0.z / 2.f);
    f3 bmax = mk(n.q0.x / 2.f, n.q0.y / 2.f, n.q0.z / 2.f);
    f3 idir = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
    float t0x = (bmin.x - o.x) * idir.x, t0y = (bmin.y - o.y) * idir.y, t0z = (bmin.z - o.z) * idir.z;
    float t1x = (bmax.x - o.x) * idir.x, t1y = (bmax.y - o.y) * idir.y, t1z = (bmax.z - o.z) * idir.z;
    float nx = fminf(t0x, t1x), ny = fminf(t0y, t1y), nz = fminf(t0z, t1z);
    float fx = fmaxf(t0x, t1x), fy = fmaxf(t0y, t1y), fz = fmaxf(t0z, t1z);
    float t_near = fmaxf(fmaxf(nx, ny), nz);
    float t_far = fminf(fminf(fx, fy), fz);
    bool along_x = d.x != 0.f && d.y == 0.f && d.z == 0.f;
    bool along_y = d.x == 0.f && d.y != 0.f && d.z == 0.f;
    bool along_z = d.x == 0.f && d.y == 0.f && d.z != 0.f;
    bool in_x = o.x > bmin.x && o.x < bmax.x;
    bool in_y = o.y > bmin.y && o.y < bmax.y;
    bool in_z = o.z > bmin.z && o.z < bmax.z;
    bool has;
    if (along_x) has = in_y && in_z;
    else if (along_y) has = in_x && in_z;
    else if (along_z) has = in_x && in_y;
    else has = (t_far > t_near && t_far > 0.f);
    if (!has) return false;
    float t = t_min < t_near ? t_near : (t_min < t_far ? t_far : t_min);
    f3 p = mk(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
    float pax = fabsf(p.x) / (bmax.x - bmin.x);
    float pay = fabsf(p.y) / (bmax.y - bmin.y);
    float paz = fabsf(p.z) / (bmax.z - bmin.z);
    f3 nn = mk(0.f, 0.f, 0.f);
    if (pax >= pay && pax >= paz) nn.x = copysignf(1.f, p.x);
    else if (pay >= pax && pay >= paz) nn.y = copysignf(1.f, p.y);
    else if (paz >= pax && paz >= pay) nn.z = copysignf(1.f, p.z);
    if (t > t_min) {
        is.x = nn.x; is.y = nn.y; is.z = nn.z; is.w = t;
        return true;
    }
    return false;
}

// csg_intersect_leaf_newcone.h:50-121
template <bool FP64>
OXD bool leaf_cone(Is& is, const NodeV& n, float t_min, f3 o, f3 d) {
    float r1 = n.q0.x, z1 = n.q0.y, r2 = n.q0.z, z2 = n.q0.w;
    const float RT_MAX = 1e16f;
    float r1r1 = r1 * r1, r2r2 = r2 * r2;
    float tth = (r2 - r1) / (z2 - z1);
    float tth2 = tth * tth;
    float z0 = (z2 * r1 - z1 * r2) / (r1 - r2);
    float idz = 1.f / d.z;
    float t_cap1 = d.z == 0.f ? RT_MAX : (z1 - o.z) * idz;
    float t_cap2 = d.z == 0.f ? RT_MAX : (z2 - o.z) * idz;
    float rr_cap1 = (o.x + t_cap1 * d.x) * (o.x + t_cap1 * d.x) + (o.y + t_cap1 * d.y) * (o.y + t_cap1 * d.y);
    float rr_cap2 = (o.x + t_cap2 * d.x) * (o.x + t_cap2 * d.x) + (o.y + t_cap2 * d.y) * (o.y + t_cap2 * d.y);
    t_cap1 = rr_cap1 < r1r1 && t_cap1 > t_min ? t_cap1 : RT_MAX;
    t_cap2 = rr_cap2 < r2r2 && t_cap2 > t_min ? t_cap2 : RT_MAX;
    float c2 = d.x * d.x + d.y * d.y - d.z * d.z * tth2;
    float c1 = o.x * d.x + o.y * d.y - (o.z - z0) * d.z * tth2;
    float c0 = o.x * o.x + o.y * o.y - (o.z - z0) * (o.z - z0) * tth2;
    float t_near, t_far, sdisc;
    if (FP64) {
        double tn, tf, sd;
        rq_roots_disq64(double(RT_MAX), tn, tf, sd, double(c2), double(c1), double(c0));
        t_near = float(tn); t_far = float(tf); sdisc = float(sd);
    } else {
        rq_roots_disq(RT_MAX, t_near, t_far, sdisc, c2, c1, c0);
    }
    float z_near = o.z + t_near * d.z;
    float z_far = o.z + t_far * d.z;
    t_near = z_near > z1 && z_near < z2 && t_near > t_min ? t_near : RT_MAX;
    t_far = z_far > z1 && z_far < z2 && t_far > t_min ? t_far : RT_MAX;
    float t = fminf(fminf(t_near, t_far), fminf(t_cap1, t_cap2));
    bool valid = t > t_min && t < RT_MAX;
    if (valid) {
        if (t == t_cap1 || t == t_cap2) {
            is.x = 0.f; is.y = 0.f; is.z = t == t_cap2 ? 1.f : -1.f;
        } else {
            f3 nn = normalize(mk(o.x + t * d.x, o.y + t * d.y, (z0 - (o.z + t * d.z)) * tth2));
            is.x = nn.x; is.y = nn.y; is.z = nn.z;
        }
        is.w = t;
    }
    return valid;
}

// csg_intersect_leaf_convexpolyhedron.h:22-106
OXD bool leaf_convexpolyhedron(Is& is, const NodeV& n, const float4* plan, float t_min, f3 o, f3 d) {
    const float INF = __int_as_float(0x7f800000);
    float t0 = -INF, t1 = INF;
    f3 t0n = mk(0.f, 0.f, 0.f), t1n = mk(0.f, 0.f, 0.f);
    unsigned planeIdx = __float_as_uint(n.q0.x), planeNum = __float_as_uint(n.q0.y);
    for (unsigned i = 0; i < planeNum; i++) {
        float4 pl = plan[planeIdx + i];
        f3 nn = mk(pl.x, pl.y, pl.z);
        float nd = dot(nn, d);
        float no = dot(nn, o);
        float dist = no - pl.w;
        float tc = -dist / nd;
        bool parallel_inside = nd == 0.f && dist < 0.f;
        bool parallel_outside = nd == 0.f && dist > 0.f;
        if (parallel_inside) continue;
        if (parallel_outside) return false;
        if (nd < 0.f) { if (tc > t0) { t0 = tc; t0n = nn; } }
        else { if (tc < t1) { t1 = tc; t1n = nn; } }
    }
    bool valid = t0 < t1;
    if (valid) {
        if (t0 > t_min) { is.x = t0n.x; is.y = t0n.y; is.z = t0n.z; is.w = t0; }
        else if (t1 > t_min) { is.x = t1n.x; is.y = t1n.y; is.z = t1n.z; is.w = t1; }
    }
    return valid;
}

// csg_intersect_leaf_phicut.h:581-635
OXD bool leaf_phicut(Is& is, const NodeV& n, float t_min, f3 o, f3 d) {
    float cosPhi0 = n.q0.x, sinPhi0 = n.q0.y, cosPhi1 = n.q0.z, sinPhi1 = n.q0.w;
    float d_n0 = d.x * sinPhi0 + d.y * (-cosPhi0);
    float d_n1 = d.x * (-sinPhi1) + d.y * (cosPhi1);
    float o_n0 = o.x * sinPhi0 + o.y * (-cosPhi0);
    float o_n1 = o.x * (-sinPhi1) + o.y * (cosPhi1);
    float t0 = d_n0 == 0.f ? t_min : -o_n0 / d_n0;
    float t1 = d_n1 == 0.f ? t_min : -o_n1 / d_n1;
    float PR = d_n0 == 0.f ? -o_n0 : -d_n0;
    float QR = d_n1 == 0.f ? -o_n1 : -d_n1;
    float PQ = cosPhi0 * sinPhi1 - cosPhi1 * sinPhi0;
    bool unbounded_exit = PQ >= 0.f ? (PR >= 0.f && QR <= 0.f) : (PR >= 0.f || QR <= 0.f);
    float side0 = o.x * cosPhi0 + o.y * sinPhi0 + (d.x * cosPhi0 + d.y * sinPhi0) * t0;
    float side1 = o.x * cosPhi1 + o.y * sinPhi1 + (d.x * cosPhi1 + d.y * sinPhi1) * t1;
    if (side0 < 0.f) t0 = t_min;
    if (side1 < 0.f) t1 = t_min;
    float t_near = fminf(t0, t1), t_far = fmaxf(t0, t1);
    float t = t_near > t_min ? t_near : (t_far > t_min ? t_far : t_min);
    bool valid = t > t_min;
    if (valid) {
        is.x = t == t1 ? -sinPhi1 : sinPhi0;
        is.y = t == t1 ? cosPhi1 : -cosPhi0;
        is.z = 0.f;
        is.w = t;
    } else if (unbounded_exit) {
        is.y = -is.y;
    }
    return valid;
}

// csg_intersect_leaf_halfspace.h:155-187: the halfspace p.n < w behind the plane p.n = w
// (q0 = unit outward normal, w). The plane crossing t = -(o.n - w)/(d.n) (t_min for a ray
// parallel to the plane) is a hit when beyond t_min. A miss whose origin lies inside
// (o.n - w < -1e-9) signals the exit at infinity with isect.y = -0 for intersect_tree.
OXD bool leaf_halfspace(Is& is, const NodeV& n, float t_min, f3 o, f3 d) {
    const f3 nn = mk(n.q0.x, n.q0.y, n.q0.z);
    const float w = n.q0.w;
    const float on = dot(o, nn), dn = dot(d, nn);
    const float on_w = on - w;
    const bool inside = on_w < -1e-9f;
    const float t = fabsf(dn) > 0.f ? -on_w / dn : t_min;
    const bool valid = t > t_min;
    if (valid) { is.x = nn.x; is.y = nn.y; is.z = nn.z; is.w = t; }
    else if (inside) { is.y = -0.f; }
    return valid;
}

// intersect_leaf (CSG/csg_intersect_leaf.h:175-324). SIMPLE compiles only the sphere,
// zsphere, cylinder and box leaves, for geometries made of nothing else: the other leaves'
// code and registers then stay out of the kernel.
template <bool FP64, bool SIMPLE = false>
OXD bool intersect_leaf(Is& is, const float4* node, int ni, const float4* plan, const float4* itra,
                        float t_min, f3 ro, f3 rd) {
    NodeV n = load_node(node, ni);
    is.x = 0.f; is.y = 0.f; is.z = 0.f; is.w = 0.f;
    unsigned tc = typecode(n);
    unsigned gtc = __float_as_uint(n.q3.w);
    unsigned gt = gtc & 0x7fffffffu;
    bool complement = (gtc & 0x80000000u) != 0;
    f3 o = ro, d = rd;
    float4 q0, q1, q2, q3;
    if (gt > 0) {
        const float4* q = itra + 4 * (gt - 1);
        q0 = q[0]; q1 = q[1]; q2 = q[2]; q3 = q[3];
        o = mk(q0.x * ro.x + q1.x * ro.y + q2.x * ro.z + q3.x * 1.f,
               q0.y * ro.x + q1.y * ro.y + q2.y * ro.z + q3.y * 1.f,
               q0.z * ro.x + q1.z * ro.y + q2.z * ro.z + q3.z * 1.f);
        d = mk(q0.x * rd.x + q1.x * rd.y + q2.x * rd.z + q3.x * 0.f,
               q0.y * rd.x + q1.y * rd.y + q2.y * rd.z + q3.y * 0.f,
               q0.z * rd.x + q1.z * rd.y + q2.z * rd.z + q3.z * 0.f);
    }
    bool valid = false;
    if constexpr (SIMPLE) {
        switch (tc) {
            case CSG_SPHERE: valid = leaf_sphere<FP64>(is, n, t_min, o, d); break;
            case CSG_ZSPHERE: valid = leaf_zsphere<FP64>(is, n, t_min, o, d); break;
            case CSG_CYLINDER: valid = leaf_cylinder<FP64>(is, n, t_min, o, d); break;
            case CSG_BOX3: valid = leaf_box3(is, n, t_min, o, d); break;
            default: break;
        }
    } else switch (tc) {
        case CSG_SPHERE: valid = leaf_sphere<FP64>(is, n, t_min, o, d); break;
        case CSG_ZSPHERE: valid = leaf_zsphere<FP64>(is, n, t_min, o, d); break;
        case CSG_CYLINDER: valid = leaf_cylinder<FP64>(is, n, t_min, o, d); break;
        case CSG_BOX3: valid = leaf_box3(is, n, t_min, o, d); break;
        case CSG_CONE: valid = leaf_cone<FP64>(is, n, t_min, o, d); break;
        case CSG_CONVEXPOLYHEDRON: valid = leaf_convexpolyhedron(is, n, plan, t_min, o, d); break;
        case CSG_PHICUT: valid = leaf_phicut(is, n, t_min, o, d); break;
        case CSG_HALFSPACE: valid = leaf_halfspace(is, n, t_min, o, d); break;
        case CSG_HYPERBOLOID: valid = leaf_hyperboloid(is, n, t_min, o, d); break;
        default: break;
    }
    if (valid && gt > 0) {
        float x = q0.x * is.x + q0.y * is.y + q0.z * is.z;
        float y = q1.x * is.x + q1.y * is.y + q1.z * is.z;
        float z = q2.x * is.x + q2.y * is.y + q2.z * is.z;
        is.x = x; is.y = y; is.z = z;
    }
    if (complement) {
        is.x = valid ? -is.x : -0.f;
        is.y = valid ? -is.y : is.y;
        is.z = valid ? -is.z : is.z;
    }
    return valid;
}

// Kensler single-hit CSG decision table, csg_classify.h:86-101
OXD int csg_lut(unsigned op, int a, int b, bool acloser) {
    const uint32_t AC = op == 1 ? 0x22121141u : (op == 2 ? 0x00014014u : 0x00141141u);
    const uint32_t BC = op == 1 ? 0x22115122u : (op == 2 ? 0x00022055u : 0x00133155u);
    unsigned offset = 3u * unsigned(a) + unsigned(b);
    uint32_t t = acloser ? AC : BC;
    return offset < 8 ? int((t >> (offset * 4)) & 0xf) : 0;
}

OXD int classify(const Is& ise, f3 dir, float tmin) {
    return fabsf(ise.w) > tmin ? ((ise.x * dir.x + ise.y * dir.y + ise.z * dir.z) < 0.f ? 0 : 1) : 2;
}

// ------------------------------------------------------------------ list nodes
// Compound nodes whose subNum leaf nodes start subOffset nodes after `root` (the prim's
// first node): CSG_CONTIGUOUS / CSG_DISCONTIGUOUS / CSG_OVERLAP (CSG/csg_intersect_node.h).
// `is` is written only for a valid intersect. Geometries without list nodes run kernels
// built with LIST=false, which contain none of this code.
constexpr int CONTIGUOUS_MAX_SUB = 8;   // enter/aux/idx arrays of csg_intersect_node.h:470-472

// intersect_node_discontiguous (csg_intersect_node.h:655-686): closest valid sub intersect
template <bool FP64>
OXD bool intersect_node_discontiguous(Is& is, const float4* node, int ni, int root, const float4* plan,
                                      const float4* itra, float t_min, f3 ro, f3 rd) {
    const unsigned num_sub = __float_as_uint(node[4 * ni].x), offset_sub = __float_as_uint(node[4 * ni].y);
    Is closest; closest.x = 0.f; closest.y = 0.f; closest.z = 0.f; closest.w = RT_DEFAULT_MAX;
    Is s; s.x = 0.f; s.y = 0.f; s.z = 0.f; s.w = 0.f;   // one per call: a leaf reporting valid without writing keeps the last values
    for (unsigned isub = 0; isub < num_sub; isub++) {
        if (intersect_leaf<FP64>(s, node, root + int(offset_sub + isub), plan, itra, t_min, ro, rd))
            if (s.w < closest.w) closest = s;
    }
    bool valid = closest.w < RT_DEFAULT_MAX;
    if (valid) is = closest;
    return valid;
}

// intersect_node_contiguous (csg_intersect_node.h:393-624): union of touching or
// overlapping subs. Zeroth pass: nearest enter and exit count; with no exits the ray
// starts outside and the nearest enter is the answer. Otherwise collect the enters,
// order them by t (insertion sort of indices), and from each enter that starts before
// the farthest exit so far, find its exit; the farthest exit is the answer.
template <bool FP64>
OXD bool intersect_node_contiguous(Is& is, const float4* node, int ni, int root, const float4* plan,
                                   const float4* itra, float t_min, f3 ro, f3 rd) {
    const int num_sub = int(__float_as_uint(node[4 * ni].x)), offset_sub = int(__float_as_uint(node[4 * ni].y));
    Is nearest_enter; nearest_enter.x = 0.f; nearest_enter.y = 0.f; nearest_enter.z = 0.f; nearest_enter.w = RT_DEFAULT_MAX;
    Is farthest_exit; farthest_exit.x = 0.f; farthest_exit.y = 0.f; farthest_exit.z = 0.f; farthest_exit.w = t_min;
    int exit_count = 0;
    const float propagate_epsilon = 0.0001f;
    Is s; s.x = 0.f; s.y = 0.f; s.z = 0.f; s.w = 0.f;   // shared by all three passes
    for (int i = 0; i < num_sub; i++) {
        if (intersect_leaf<FP64>(s, node, root + offset_sub + i, plan, itra, t_min, ro, rd)) {
            int st = classify(s, rd, t_min);
            if (st == 0) { if (s.w < nearest_enter.w) nearest_enter = s; }
            else if (st == 1) exit_count += 1;
        }
    }
    if (exit_count == 0) {
        bool valid = nearest_enter.w > t_min && nearest_enter.w < RT_DEFAULT_MAX;
        if (valid) is = nearest_enter;
        return valid;
    }
    int enter_count = 0;
    float enter[CONTIGUOUS_MAX_SUB];
    int aux[CONTIGUOUS_MAX_SUB], idx[CONTIGUOUS_MAX_SUB];
    for (int isub = 0; isub < num_sub; isub++) {
        if (intersect_leaf<FP64>(s, node, root + offset_sub + isub, plan, itra, t_min, ro, rd)) {
            int st = classify(s, rd, t_min);
            if (st == 0) {
                aux[enter_count] = isub; idx[enter_count] = enter_count; enter[enter_count] = s.w;
                enter_count += 1;
            } else if (st == 1) {
                exit_count += 1;
                if (s.w > farthest_exit.w) farthest_exit = s;
            }
        }
    }
    for (int i = 1; i < enter_count; i++) {
        int key = idx[i];
        int j = i - 1;
        while (j >= 0 && enter[idx[j]] > enter[key]) { idx[j + 1] = idx[j]; j = j - 1; }
        idx[j + 1] = key;
    }
    for (int i = 0; i < enter_count; i++) {
        float tminAdvanced = enter[idx[i]] + propagate_epsilon;
        int isub = aux[idx[i]];
        if (tminAdvanced < farthest_exit.w) {
            if (intersect_leaf<FP64>(s, node, root + offset_sub + isub, plan, itra, tminAdvanced, ro, rd)) {
                if (classify(s, rd, tminAdvanced) == 1) {
                    exit_count += 1;
                    if (s.w > farthest_exit.w) farthest_exit = s;
                }
            }
        }
    }
    bool valid = exit_count > 0 && farthest_exit.w > t_min;
    if (valid) is = farthest_exit;
    return valid;
}

// intersect_node_overlap (csg_intersect_node.h:820-895): intersection of all subs; the
// farthest enter (or, from inside, the nearest exit) when every sub takes part
template <bool FP64>
OXD bool intersect_node_overlap(Is& is, const float4* node, int ni, int root, const float4* plan,
                                const float4* itra, float t_min, f3 ro, f3 rd) {
    const unsigned num_sub = __float_as_uint(node[4 * ni].x), offset_sub = __float_as_uint(node[4 * ni].y);
    Is farthest_enter; farthest_enter.x = 0.f; farthest_enter.y = 0.f; farthest_enter.z = 0.f; farthest_enter.w = t_min;
    Is nearest_exit; nearest_exit.x = 0.f; nearest_exit.y = 0.f; nearest_exit.z = 0.f; nearest_exit.w = RT_DEFAULT_MAX;
    const float propagate_epsilon = 0.0001f;
    unsigned enter_count = 0, exit_count = 0;
    Is s; s.x = 0.f; s.y = 0.f; s.z = 0.f; s.w = 0.f;   // one per call, reused for the exit search
    for (unsigned isub = 0; isub < num_sub; isub++) {
        const int sn = root + int(offset_sub + isub);
        if (intersect_leaf<FP64>(s, node, sn, plan, itra, t_min, ro, rd)) {
            int st = classify(s, rd, t_min);
            if (st == 0) {
                enter_count += 1;
                if (s.w > farthest_enter.w) farthest_enter = s;
                float tminAdvanced = s.w + propagate_epsilon;
                if (intersect_leaf<FP64>(s, node, sn, plan, itra, tminAdvanced, ro, rd)) {
                    if (classify(s, rd, tminAdvanced) == 1) {
                        exit_count += 1;
                        if (s.w < nearest_exit.w) nearest_exit = s;
                    }
                }
            } else if (st == 1) {
                exit_count += 1;
                if (s.w < nearest_exit.w) nearest_exit = s;
            }
        }
    }
    bool overlap_all = farthest_enter.w < nearest_exit.w && (enter_count > exit_count ? enter_count : exit_count) == num_sub;
    if (overlap_all) {
        if (farthest_enter.w > t_min && farthest_enter.w < RT_DEFAULT_MAX) { is = farthest_enter; return true; }
        if (nearest_exit.w > t_min && nearest_exit.w < RT_DEFAULT_MAX) { is = nearest_exit; return true; }
    }
    return false;
}

// intersect_node (csg_intersect_node.h:914-938): list node or leaf
template <bool FP64, bool LIST, bool SIMPLE = false>
OXD bool intersect_node(Is& is, const float4* node, int ni, int root, const float4* plan, const float4* itra,
                        float t_min, f3 ro, f3 rd) {
    if (!LIST) return intersect_leaf<FP64, SIMPLE>(is, node, ni, plan, itra, t_min, ro, rd);
    const unsigned tc = __float_as_uint(node[4 * ni + 3].z);
    if (tc == CSG_CONTIGUOUS) { is.x = 0.f; is.y = 0.f; is.z = 0.f; is.w = 0.f; return intersect_node_contiguous<FP64>(is, node, ni, root, plan, itra, t_min, ro, rd); }
    if (tc == CSG_DISCONTIGUOUS) { is.x = 0.f; is.y = 0.f; is.z = 0.f; is.w = 0.f; return intersect_node_discontiguous<FP64>(is, node, ni, root, plan, itra, t_min, ro, rd); }
    if (tc == CSG_OVERLAP) { is.x = 0.f; is.y = 0.f; is.z = 0.f; is.w = 0.f; return intersect_node_overlap<FP64>(is, node, ni, root, plan, itra, t_min, ro, rd); }
    return intersect_leaf<FP64>(is, node, ni, plan, itra, t_min, ro, rd);
}

// intersect_tree (CSG/csg_intersect_tree.h:277-672): postorder evaluation with a tranche
// stack of subtree slices and a stack of partial intersects; side recorded in sign(t).
// The general form keeps the reference's capacities, 15 partial intersects and 4
// tranches, in private arrays. Over every sequence of LUT decisions a tree of height h
// holds at most h+1 partial intersects, and a tree of height <= 2 at most 3 tranches:
// with SHALLOW (every tree of the geometry has height <= 2) both stacks are three named
// local entries each, selected by if-chains and never indexed, so they live in registers
// and no private memory is touched. The overflow checks keep the reference's meaning:
// with these capacities they cannot fire for such trees, as they cannot with its own.
template <bool FP64, bool LIST, bool SHALLOW = false, bool SIMPLE = false>
OXD bool intersect_tree(Is& out, const float4* node, int root, const float4* plan, const float4* itra,
                        float t_min, f3 ro, f3 rd) {
    const unsigned numNode = __float_as_uint(node[4 * root].x);
    const unsigned height = unsigned(__ffs(int(numNode + 1)) - 2);
    const float propagate_epsilon = 0.0001f;
    constexpr int NS = SHALLOW ? 3 : 15, NT = SHALLOW ? 3 : 4;
    Is st[SHALLOW ? 1 : 15];                       // general form
    unsigned tr_slice[SHALLOW ? 1 : 4];
    float tr_tmin[SHALLOW ? 1 : 4];
    Is s0, s1, s2;                                 // SHALLOW form
    s0.x = s0.y = s0.z = s0.w = 0.f; s1 = s0; s2 = s0;
    unsigned c0 = 0u, c1 = 0u, c2 = 0u;
    float m0 = 0.f, m1 = 0.f, m2 = 0.f;
    // SHALLOW accessors: every entry is written on every update (a select per component),
    // so no store depends on control flow and no entry ever needs an address
    auto put = [&](int i, const Is& v) __attribute__((always_inline)) {
        if constexpr (SHALLOW) {
            const bool k0 = i == 0, k1 = i == 1, k2 = i >= 2;
            s0.x = k0 ? v.x : s0.x; s0.y = k0 ? v.y : s0.y; s0.z = k0 ? v.z : s0.z; s0.w = k0 ? v.w : s0.w;
            s1.x = k1 ? v.x : s1.x; s1.y = k1 ? v.y : s1.y; s1.z = k1 ? v.z : s1.z; s1.w = k1 ? v.w : s1.w;
            s2.x = k2 ? v.x : s2.x; s2.y = k2 ? v.y : s2.y; s2.z = k2 ? v.z : s2.z; s2.w = k2 ? v.w : s2.w;
        } else {
            st[i] = v;
        }
    };
    auto take = [&](int i) __attribute__((always_inline)) -> Is {
        Is r;
        if constexpr (SHALLOW) {
            const float x01 = i == 0 ? s0.x : s1.x, y01 = i == 0 ? s0.y : s1.y;
            const float z01 = i == 0 ? s0.z : s1.z, w01 = i == 0 ? s0.w : s1.w;
            r.x = i >= 2 ? s2.x : x01; r.y = i >= 2 ? s2.y : y01; r.z = i >= 2 ? s2.z : z01; r.w = i >= 2 ? s2.w : w01;
        } else {
            r = st[i];
        }
        return r;
    };
    auto push_tr = [&](int i, unsigned c, float m) __attribute__((always_inline)) {
        if constexpr (SHALLOW) {
            const bool k0 = i == 0, k1 = i == 1, k2 = i >= 2;
            c0 = k0 ? c : c0; m0 = k0 ? m : m0;
            c1 = k1 ? c : c1; m1 = k1 ? m : m1;
            c2 = k2 ? c : c2; m2 = k2 ? m : m2;
        } else {
            tr_slice[i] = c; tr_tmin[i] = m;
        }
    };
    auto pick = [](bool c, const Is& a, const Is& b) __attribute__((always_inline)) -> Is {
        Is r;
        r.x = c ? a.x : b.x; r.y = c ? a.y : b.y; r.z = c ? a.z : b.z; r.w = c ? a.w : b.w;
        return r;
    };
    int tr = -1;
    int sp = -1;
    int ierr = 0;
    tr++; push_tr(tr, ((1u << height) & 0xffu) << 16, t_min);
    while (tr > -1) {
        unsigned slice;
        float tmin;
        if constexpr (SHALLOW) {
            slice = tr >= 2 ? c2 : (tr == 1 ? c1 : c0);
            tmin = tr >= 2 ? m2 : (tr == 1 ? m1 : m0);
        } else {
            slice = tr_slice[tr];
            tmin = tr_tmin[tr];
        }
        tr--;
        unsigned nodeIdx = (slice >> 16) & 0xff, endIdx = (slice >> 24) & 0xff;
        while (nodeIdx != endIdx) {
            unsigned depth = 31u - unsigned(__clz(int(nodeIdx)));
            unsigned elevation = height - depth;
            int ni = root + int(nodeIdx) - 1;
            unsigned tc = __float_as_uint(node[4 * ni + 3].z);
            if (tc == CSG_ZERO) {
                nodeIdx = (nodeIdx & 1u) ? nodeIdx >> 1 : (nodeIdx << elevation) + (1u << elevation);
                continue;
            }
            if (tc >= CSG_NODE) {
                Is ni_is; ni_is.x = 0.f; ni_is.y = 0.f; ni_is.z = 0.f; ni_is.w = 0.f;
                intersect_node<FP64, LIST, SIMPLE>(ni_is, node, ni, root, plan, itra, tmin, ro, rd);
                ni_is.w = copysignf(ni_is.w, nodeIdx % 2 == 0 ? -1.f : 1.f);
                if (sp >= NS - 1) { ierr = 1; break; }
                put(++sp, ni_is);
            } else {
                if (sp < 1) { ierr = 1; break; }
                const Is top0 = take(sp), top1 = take(sp - 1);
                bool firstLeft = signbit(top0.w);
                bool secondLeft = signbit(top1.w);
                if (!(firstLeft ^ secondLeft)) { ierr = 1; break; }
                const Is L = pick(firstLeft, top0, top1), R = pick(firstLeft, top1, top0);
                int l_state = classify(L, rd, tmin);
                int r_state = classify(R, rd, tmin);
                float t_left = fabsf(L.w), t_right = fabsf(R.w);
                bool leftIsCloser = t_left <= t_right;
                bool l_promote = l_state == 2 && (signbit(L.x) || signbit(L.y));
                bool r_promote = r_state == 2 && (signbit(R.x) || signbit(R.y));
                if (r_promote) { r_state = 1; leftIsCloser = true; }
                if (l_promote) { l_state = 1; leftIsCloser = false; }
                int ctrl = csg_lut(tc, l_state, r_state, leftIsCloser);
                if (ctrl < 4) {
                    Is result = pick(ctrl == 1, L, R);
                    if (ctrl == 0) { result.x = 0.f; result.y = 0.f; result.z = 0.f; result.w = 0.f; }
                    if (ctrl == 3) { result.x = -result.x; result.y = -result.y; result.z = -result.z; }
                    result.w = copysignf(result.w, nodeIdx % 2 == 0 ? -1.f : 1.f);
                    sp -= 2;
                    put(++sp, result);
                } else {
                    bool loopA = ctrl == 4;
                    const Is other = pick(loopA, R, L);
                    unsigned leftIdx = 2 * nodeIdx, rightIdx = leftIdx + 1;
                    float tminAdvanced = fabsf(loopA ? L.w : R.w) + propagate_epsilon;
                    sp -= 2;
                    put(++sp, other);
                    unsigned endTree = ((nodeIdx & 0xffu) << 16) | ((endIdx & 0xffu) << 24);
                    unsigned leftTree = (((leftIdx << (elevation - 1)) & 0xffu) << 16) | (((rightIdx << (elevation - 1)) & 0xffu) << 24);
                    unsigned rightTree = (((rightIdx << (elevation - 1)) & 0xffu) << 16) | ((nodeIdx & 0xffu) << 24);
                    unsigned loopTree = loopA ? leftTree : rightTree;
                    if (tr >= NT - 1) { ierr = 1; break; }
                    tr++; push_tr(tr, endTree, tmin);
                    if (tr >= NT - 1) { ierr = 1; break; }
                    tr++; push_tr(tr, loopTree, tminAdvanced);
                    break;
                }
            }
            nodeIdx = (nodeIdx & 1u) ? nodeIdx >> 1 : (nodeIdx << elevation) + (1u << elevation);
        }
        if (ierr) break;
    }
    out.x = 0.f; out.y = 0.f; out.z = 0.f; out.w = 0.f;
    if (sp == 0) out = take(0);
    return out.w > 0.f;
}

// intersect_prim (CSG/csg_intersect_tree.h:684-719)
template <bool FP64, bool LIST, bool SHALLOW = false, bool SIMPLE = false>
OXD bool intersect_prim(Is& is, const float4* node, int root, const float4* plan, const float4* itra,
                        float t_min, f3 ro, f3 rd) {
    is.x = 0.f; is.y = 0.f; is.z = 0.f; is.w = 0.f;   // zeroed isect of CSGOptiX7.cu:894-898
    unsigned tc = __float_as_uint(node[4 * root + 3].z);
    if (tc >= CSG_LEAF) return intersect_leaf<FP64, SIMPLE>(is, node, root, plan, itra, t_min, ro, rd);
    if (tc < CSG_NODE) return intersect_tree<FP64, LIST, SHALLOW, SIMPLE>(is, node, root, plan, itra, t_min, ro, rd);
    if (LIST) return intersect_node<FP64, LIST>(is, node, root, root, plan, itra, t_min, ro, rd);
    is.x = 0.f; is.y = 0.f; is.z = 0.f; is.w = 0.f;
    return false;
}

} // namespace oxd
