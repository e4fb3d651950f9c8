// oracle.cpp -- CPU restatement of the eic-opticks simulate path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP engine: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The product library
// (eic-opticks_amd/csrc) never includes, links or calls it.
//
// Templated on the arithmetic type T: T=float follows the reference's FP32 device code
// operation by operation (same expression order, no FMA contraction, correctly rounded
// sqrt/div, transcendental functions evaluated in double and rounded), T=double is the
// FP64 oracle for position/time tolerances.
//
// Each function cites the reference file:line it restates. Parity pins (what checks
// each part independently of the GPU) are listed in tests/test_oracle_*.py.
#include <cstdint>
#include <cstring>
#include <cmath>
#include <vector>
#include <thread>
#include <atomic>
#include <algorithm>
#include <limits>
#include "../include/oxr.h"

namespace orc {

// ---------------------------------------------------------------- small vectors
template <typename T> struct V3 { T x, y, z; };
template <typename T> struct V4 { T x, y, z, w; };
template <typename T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <typename T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <typename T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <typename T> inline V3<T> normalize(V3<T> a) { T inv = T(1) / std::sqrt(dot(a, a)); return a * inv; }

// transcendental functions: double evaluation rounded to T
template <typename T> inline T xlog(T x) { return T(std::log(double(x))); }
// single precision: the device's log (FreeBSD msun e_logf.c reduction and polynomial in
// IEEE single operations), restated so both give the same bits
template <> inline float xlog<float>(float x) {
    uint32_t ix; std::memcpy(&ix, &x, 4);
    if (ix >= 0x7f800000u || ix < 0x00800000u) return float(std::log(double(x)));
    const float ln2_hi = 6.9313812256e-01f, ln2_lo = 9.0580006145e-06f;
    const float Lg1 = 0.66666662693f, Lg2 = 0.40000972152f, Lg3 = 0.28498786688f, Lg4 = 0.24279078841f;
    int k = int(ix >> 23) - 127;
    uint32_t m = ix & 0x007fffffu;
    uint32_t i = (m + 0x004afb20u) & 0x00800000u;
    uint32_t xb = m | (i ^ 0x3f800000u);
    float xm; std::memcpy(&xm, &xb, 4);
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
template <typename T> inline T xsin(T x) { return T(std::sin(double(x))); }
template <typename T> inline T xcos(T x) { return T(std::cos(double(x))); }
template <typename T> inline T xatan2(T y, T x) { return T(std::atan2(double(y), double(x))); }
template <typename T> inline T pi() { return T(3.14159265358979323846); }

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// ---------------------------------------------------------------- flags (sysrap/OpticksPhoton.h:22-46)
enum : uint32_t {
    CERENKOV = 1u << 0, SCINTILLATION = 1u << 1, TORCH = 1u << 2, BULK_ABSORB = 1u << 3,
    BULK_REEMIT = 1u << 4, BULK_SCATTER = 1u << 5, SURFACE_DETECT = 1u << 6,
    SURFACE_ABSORB = 1u << 7, SURFACE_DREFLECT = 1u << 8, SURFACE_SREFLECT = 1u << 9,
    BOUNDARY_REFLECT = 1u << 10, BOUNDARY_TRANSMIT = 1u << 11, MISS = 1u << 15
};
enum { UNDEFINED = 0, BREAK = 1, CONTINUE = 2, BOUNDARY = 3 };   // sysrap/sflow.h:10-18
enum { CSG_ZERO = 0, CSG_UNION = 1, CSG_INTERSECTION = 2, CSG_DIFFERENCE = 3, CSG_NODE = 11,
       CSG_CONTIGUOUS = 11, CSG_DISCONTIGUOUS = 12, CSG_OVERLAP = 13, CSG_LEAF = 101,
       CSG_SPHERE = 101, CSG_ZSPHERE = 103, CSG_CYLINDER = 105, CSG_CONE = 108, CSG_BOX3 = 110,
       CSG_CONVEXPOLYHEDRON = 112, CSG_HYPERBOLOID = 117, CSG_PHICUT = 121, CSG_HALFSPACE = 125 };

// ---------------------------------------------------------------- Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11:
// 10 rounds of the 4x32 Philox bijection with multipliers 0xD2511F53, 0xCD9E8D57 and
// Weyl key increments 0x9E3779B9, 0xBB67AE85.
inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Per-photon uniform stream. key = seed, counter = (block lo, block hi, photon lo, photon hi);
// the stream of photon p starts at uniform index offset + skipahead*event (qrng.h:131-137
// semantics: subsequence = photon index, skipahead per event). u = ((x >> 8) + 1) * 2^-24,
// which lies in (0,1] with 24 significant bits.
struct Stream {
    uint32_t key[2];
    uint64_t photon;
    uint64_t pos;            // absolute uniform index
    const float* pre;        // precooked row, or null
    int32_t npre;
    int32_t cursor;
    uint32_t overflow;
    uint32_t buf[4];
    uint64_t buf_block;
    void init(uint64_t seed, uint64_t offset, uint64_t skip, uint64_t event, uint64_t photon_idx) {
        key[0] = uint32_t(seed); key[1] = uint32_t(seed >> 32);
        photon = photon_idx;
        pos = offset + skip * event;
        pre = nullptr; npre = 0; cursor = 0; overflow = 0;
        buf_block = ~0ull;
    }
    float next() {
        if (pre) {
            if (cursor >= npre) { cursor = 0; overflow++; }
            return pre[cursor++];
        }
        uint64_t block = pos >> 2;
        if (block != buf_block) {
            uint32_t ctr[4] = {uint32_t(block), uint32_t(block >> 32), uint32_t(photon), uint32_t(photon >> 32)};
            philox4x32_10(ctr, key, buf);
            buf_block = block;
        }
        uint32_t x = buf[pos & 3];
        pos++;
        return float((x >> 8) + 1u) * (1.0f / 16777216.0f);
    }
};

// ---------------------------------------------------------------- data views
struct Node { float f[16]; uint32_t u(int i) const { return f2u(f[i]); } };
struct Inst { float f[16]; };

template <typename T> struct Q4x3 { T m[4][3]; };   // row-convention affine (rows 0-2 linear, row 3 translation)

struct Ctx {
    // geometry
    const int32_t* solid; int64_t num_solid;
    const float* prim; int64_t num_prim;
    const Node* node; int64_t num_node;
    const float* plan; int64_t num_plan;
    const float* itra; int64_t num_itra;
    const float* inst; int64_t num_inst;
    std::vector<Q4x3<double>> w2o;    // per-instance world-to-object (double)
    std::vector<uint8_t> inst_identity_transform;
    // optional conservative culling (world box per instance, object box per prim); off by
    // default, validated against the unculled loop in tests
    int cull = 0;
    std::vector<double> inst_box;     // 6 per instance
    std::vector<double> prim_box;     // 6 per prim
    // physics
    const float* bnd; int32_t num_bnd, num_wl; float wl_min, wl_step;
    const uint32_t* optical;
    const float* icdf; int32_t icdf_rows, icdf_width, hd_factor;
    oxr_config cfg;
};

// ---------------------------------------------------------------- tables
// qbnd::boundary_lookup (qudarap/qbnd.h:103-125): linear-filtered, normalised-coordinate,
// wrap-addressed float4 texture of width num_wl and height 2*4*num_bnd. The texel
// coordinate is formed exactly as the texture unit would: x*(width) - 0.5.
template <typename T>
inline V4<T> boundary_lookup(const Ctx& c, T nm, unsigned line, unsigned k) {
    const int nx = c.num_wl;
    T fx = (nm - T(c.wl_min)) / T(c.wl_step);
    T x = (fx + T(0.5)) / T(nx);
    unsigned iy = 2u * line + k;
    T xf = x * T(nx) - T(0.5);
    T fl = std::floor(xf);
    T a = xf - fl;
    long i0 = long(fl);
    long m0 = ((i0 % nx) + nx) % nx;
    long m1 = (m0 + 1) % nx;
    const float* r = c.bnd + size_t(iy) * nx * 4;
    const float* p0 = r + m0 * 4;
    const float* p1 = r + m1 * 4;
    T b = T(1) - a;
    return V4<T>{b * T(p0[0]) + a * T(p1[0]), b * T(p0[1]) + a * T(p1[1]),
                 b * T(p0[2]) + a * T(p1[2]), b * T(p0[3]) + a * T(p1[3])};
}

// qscint tex2D(scint_tex, x, y) with y at a texel centre (qscint.h:160-228)
template <typename T>
inline T icdf_lookup(const Ctx& c, T x, int row) {
    const int nx = c.icdf_width;
    T xf = x * T(nx) - T(0.5);
    T fl = std::floor(xf);
    T a = xf - fl;
    long i0 = long(fl);
    long m0 = ((i0 % nx) + nx) % nx;
    long m1 = (m0 + 1) % nx;
    const float* r = c.icdf + size_t(row) * nx;
    return (T(1) - a) * T(r[m0]) + a * T(r[m1]);
}

// qscint::wavelength_hd0/10/20 (qudarap/qscint.h:150-228)
template <typename T>
inline T scint_wavelength(const Ctx& c, T u0) {
    if (!c.icdf) return T(0);
    switch (c.hd_factor) {
        case 0: return icdf_lookup<T>(c, u0, 0);
        case 10:
            if (u0 < T(0.1)) return icdf_lookup<T>(c, u0 * T(10), 1);
            if (u0 > T(0.9)) return icdf_lookup<T>(c, (u0 - T(0.9)) * T(10), 2);
            return icdf_lookup<T>(c, u0, 0);
        case 20:
            if (u0 < T(0.05)) return icdf_lookup<T>(c, u0 * T(20), 1);
            if (u0 > T(0.95)) return icdf_lookup<T>(c, (u0 - T(0.95)) * T(20), 2);
            return icdf_lookup<T>(c, u0, 0);
    }
    return T(0);
}

// ---------------------------------------------------------------- CSG leaves
template <typename T> struct Isect { T x, y, z, w; };

// csg_robust_quadratic_roots.h:95-111 : d t^2 + 2 b t + c = 0
template <typename T>
inline void robust_quadratic_roots(T& t1, T& t2, T& disc, T& sdisc, T d, T b, T c) {
    disc = b * b - d * c;
    sdisc = disc > T(0) ? std::sqrt(disc) : T(0);
    T q = b > T(0) ? -(b + sdisc) : -(b - sdisc);
    T root1 = q / d;
    T root2 = c / q;
    t1 = std::fmin(root1, root2);   // fminf/fmaxf: a NaN root yields the other
    t2 = std::fmax(root1, root2);
}
// csg_robust_quadratic_roots.h:125-134
template <typename T>
inline void robust_quadratic_roots_disqualifying(T t_min, T& t1, T& t2, T& disc, T& sdisc, T d, T b, T c) {
    disc = b * b - d * c;
    sdisc = disc > T(0) ? std::sqrt(disc) : T(0);
    T q = b > T(0) ? -(b + sdisc) : -(b - sdisc);
    T root1 = sdisc > T(0) ? q / d : t_min;
    T root2 = sdisc > T(0) ? c / q : t_min;
    t1 = std::fmin(root1, root2);   // fminf/fmaxf: a NaN root yields the other
    t2 = std::fmax(root1, root2);
}

// fminf/fmaxf semantics (NaN-ignoring) for the cases that rely on them
template <typename T> inline T fmn(T a, T b) { return std::fmin(a, b); }
template <typename T> inline T fmx(T a, T b) { return std::fmax(a, b); }

// csg_intersect_leaf_sphere.h:15-55 ; fp64 computes b, c, d and roots in double.
template <typename T>
void leaf_sphere(bool& valid, Isect<T>& is, const Node& n, T t_min, V3<T> o, V3<T> dd, bool fp64) {
    V3<T> center = v3<T>(n.f[0], n.f[1], n.f[2]);
    T radius = n.f[3];
    V3<T> O = o - center;
    V3<T> D = dd;
    T root1, root2, disc, sdisc;
    if (fp64) {
        V3<double> Od = v3<double>(O.x, O.y, O.z), Dd = v3<double>(D.x, D.y, D.z);
        double b = dot(Od, Dd), c = dot(Od, Od) - double(radius) * double(radius), d = dot(Dd, Dd);
        double r1, r2, di, sd;
        robust_quadratic_roots<double>(r1, r2, di, sd, d, b, c);
        root1 = T(r1); root2 = T(r2); disc = T(di); sdisc = T(sd);
    } else {
        T b = dot(O, D);
        T c = dot(O, O) - radius * radius;
        T d = dot(D, D);
        robust_quadratic_roots(root1, root2, disc, sdisc, d, b, c);
    }
    T t_cand = sdisc > T(0) ? (root1 > t_min ? root1 : root2) : t_min;
    valid = t_cand > t_min;
    if (valid) {
        is.x = (O.x + t_cand * D.x) / radius;
        is.y = (O.y + t_cand * D.y) / radius;
        is.z = (O.z + t_cand * D.z) / radius;
        is.w = t_cand;
    }
}

// csg_intersect_leaf_zsphere.h:45-151
template <typename T>
void leaf_zsphere(bool& valid, Isect<T>& is, const Node& n, T t_min, V3<T> ray_origin, V3<T> ray_direction, bool fp64) {
    V3<T> center = v3<T>(n.f[0], n.f[1], n.f[2]);
    V3<T> O = ray_origin - center;
    V3<T> D = ray_direction;
    T radius = n.f[3];
    T b = dot(O, D);
    T c = dot(O, O) - radius * radius;
    if (c > T(0) && b > T(0)) { valid = false; return; }
    T zmax = center.z + T(n.f[5]);
    T zmin = center.z + T(n.f[4]);
    T d = dot(D, D);
    T t1sph, t2sph, disc, sdisc;
    if (fp64) {
        V3<double> Od = v3<double>(O.x, O.y, O.z), Dd = v3<double>(D.x, D.y, D.z);
        double bd = dot(Od, Dd), cd = dot(Od, Od) - double(radius) * double(radius), dd = dot(Dd, Dd);
        double r1, r2, di, sd;
        robust_quadratic_roots<double>(r1, r2, di, sd, dd, bd, cd);
        t1sph = T(r1); t2sph = T(r2); disc = T(di); sdisc = T(sd);
    } else {
        robust_quadratic_roots(t1sph, t2sph, disc, sdisc, d, b, c);
    }
    T z1sph = ray_origin.z + t1sph * ray_direction.z;
    T z2sph = ray_origin.z + t2sph * ray_direction.z;
    T idz = T(1) / ray_direction.z;
    T t_QCAP = (zmax - ray_origin.z) * idz;
    T t_PCAP = (zmin - ray_origin.z) * idz;
    T t1cap = fmn(t_QCAP, t_PCAP);
    T t2cap = fmx(t_QCAP, t_PCAP);
    if (t1cap < t1sph || t1cap > t2sph) t1cap = t_min;
    if (t2cap < t1sph || t2cap > t2sph) t2cap = t_min;
    T t_cand = t_min;
    if (sdisc > T(0)) {
        if (t1sph > t_min && z1sph > zmin && z1sph <= zmax) t_cand = t1sph;
        else if (t1cap > t_min) t_cand = t1cap;
        else if (t2cap > t_min) t_cand = t2cap;
        else if (t2sph > t_min && z2sph > zmin && z2sph <= zmax) t_cand = t2sph;
    }
    valid = t_cand > t_min;
    if (valid) {
        is.w = t_cand;
        if (t_cand == t1sph || t_cand == t2sph) {
            is.x = (O.x + t_cand * D.x) / radius;
            is.y = (O.y + t_cand * D.y) / radius;
            is.z = (O.z + t_cand * D.z) / radius;
        } else {
            is.x = T(0); is.y = T(0);
            is.z = t_cand == t_PCAP ? T(-1) : T(1);
        }
    }
}

// csg_intersect_leaf_cylinder.h:34-82
template <typename T>
void leaf_cylinder(bool& valid, Isect<T>& is, const Node& n, T t_min, V3<T> o, V3<T> v, bool fp64) {
    T r = n.f[3], z1 = n.f[4], z2 = n.f[5];
    T ox = o.x, oy = o.y, oz = o.z, vx = v.x, vy = v.y, vz = v.z;
    T r2 = r * r;
    T a = vx * vx + vy * vy;
    T b = ox * vx + oy * vy;
    T c = ox * ox + oy * oy - r2;
    T t_near, t_far, disc, sdisc;
    if (fp64) {
        double ad = double(vx) * vx + double(vy) * vy, bd = double(ox) * vx + double(oy) * vy;
        double cd = double(ox) * ox + double(oy) * oy - double(r) * double(r);
        double tn, tf, di, sd;
        robust_quadratic_roots_disqualifying<double>(double(t_min), tn, tf, di, sd, ad, bd, cd);
        t_near = T(tn); t_far = T(tf); disc = T(di); sdisc = T(sd);
    } else {
        robust_quadratic_roots_disqualifying(t_min, t_near, t_far, disc, sdisc, a, b, c);
    }
    T z_near = oz + t_near * vz;
    T z_far = oz + t_far * vz;
    T t_z1cap = (z1 - oz) / vz;
    T r2_z1cap = (ox + t_z1cap * vx) * (ox + t_z1cap * vx) + (oy + t_z1cap * vy) * (oy + t_z1cap * vy);
    T t_z2cap = (z2 - oz) / vz;
    T r2_z2cap = (ox + t_z2cap * vx) * (ox + t_z2cap * vx) + (oy + t_z2cap * vy) * (oy + t_z2cap * vy);
    const T INF = std::numeric_limits<T>::infinity();
    T t_cand = INF;
    if (t_near > t_min && z_near > z1 && z_near < z2 && t_near < t_cand) t_cand = t_near;
    if (t_far > t_min && z_far > z1 && z_far < z2 && t_far < t_cand) t_cand = t_far;
    if (t_z1cap > t_min && r2_z1cap <= r2 && t_z1cap < t_cand) t_cand = t_z1cap;
    if (t_z2cap > t_min && r2_z2cap <= r2 && t_z2cap < t_cand) t_cand = t_z2cap;
    valid = t_cand > t_min && t_cand < INF;
    if (valid) {
        bool sheet = (t_cand == t_near || t_cand == t_far);
        is.x = sheet ? (ox + t_cand * vx) / r : T(0);
        is.y = sheet ? (oy + t_cand * vy) / r : T(0);
        is.z = sheet ? T(0) : (t_cand == t_z1cap ? T(-1) : T(1));
        is.w = t_cand;
    }
}

// csg_intersect_leaf_box3.h:64-157
template <typename T>
void leaf_box3(bool& valid, Isect<T>& is, const Node& n, T t_min, V3<T> o, V3<T> d) {
    V3<T> bmin = v3<T>(-T(n.f[0]) / T(2), -T(n.f[1]) / T(2), -T(n.f[2]) / T(2));
    V3<T> bmax = v3<T>(T(n.f[0]) / T(2), T(n.f[1]) / T(2), T(n.f[2]) / T(2));
    V3<T> idir = v3<T>(T(1) / d.x, T(1) / d.y, T(1) / d.z);
    V3<T> t0 = v3<T>((bmin.x - o.x) * idir.x, (bmin.y - o.y) * idir.y, (bmin.z - o.z) * idir.z);
    V3<T> t1 = v3<T>((bmax.x - o.x) * idir.x, (bmax.y - o.y) * idir.y, (bmax.z - o.z) * idir.z);
    V3<T> nr = v3<T>(fmn(t0.x, t1.x), fmn(t0.y, t1.y), fmn(t0.z, t1.z));
    V3<T> fr = v3<T>(fmx(t0.x, t1.x), fmx(t0.y, t1.y), fmx(t0.z, t1.z));
    T t_near = fmx(fmx(nr.x, nr.y), nr.z);
    T t_far = fmn(fmn(fr.x, fr.y), fr.z);
    bool along_x = d.x != T(0) && d.y == T(0) && d.z == T(0);
    bool along_y = d.x == T(0) && d.y != T(0) && d.z == T(0);
    bool along_z = d.x == T(0) && d.y == T(0) && d.z != T(0);
    bool in_x = o.x > bmin.x && o.x < bmax.x;
    bool in_y = o.y > bmin.y && o.y < bmax.y;
    bool in_z = o.z > bmin.z && o.z < bmax.z;
    bool has;
    if (along_x) has = in_y && in_z;
    else if (along_y) has = in_x && in_z;
    else if (along_z) has = in_x && in_y;
    else has = (t_far > t_near && t_far > T(0));
    valid = false;
    if (has) {
        T t_cand = t_min < t_near ? t_near : (t_min < t_far ? t_far : t_min);
        V3<T> p = v3<T>(o.x + t_cand * d.x, o.y + t_cand * d.y, o.z + t_cand * d.z);
        V3<T> pa = v3<T>(std::fabs(p.x) / (bmax.x - bmin.x), std::fabs(p.y) / (bmax.y - bmin.y),
                         std::fabs(p.z) / (bmax.z - bmin.z));
        V3<T> nn = v3<T>(0, 0, 0);
        if (pa.x >= pa.y && pa.x >= pa.z) nn.x = std::copysign(T(1), p.x);
        else if (pa.y >= pa.x && pa.y >= pa.z) nn.y = std::copysign(T(1), p.y);
        else if (pa.z >= pa.x && pa.z >= pa.y) nn.z = std::copysign(T(1), p.z);
        if (t_cand > t_min) {
            valid = true;
            is.x = nn.x; is.y = nn.y; is.z = nn.z; is.w = t_cand;
        }
    }
}

// csg_intersect_leaf_newcone.h:50-121
template <typename T>
void leaf_cone(bool& valid, Isect<T>& is, const Node& n, T t_min, V3<T> o, V3<T> d, bool fp64) {
    T r1 = n.f[0], z1 = n.f[1], r2 = n.f[2], z2 = n.f[3];
    const T RT_MAX = T(1e16f);
    T r1r1 = r1 * r1, r2r2 = r2 * r2;
    T tth = (r2 - r1) / (z2 - z1);
    T tth2 = tth * tth;
    T z0 = (z2 * r1 - z1 * r2) / (r1 - r2);
    T idz = T(1) / d.z;
    T t_cap1 = d.z == T(0) ? RT_MAX : (z1 - o.z) * idz;
    T t_cap2 = d.z == T(0) ? RT_MAX : (z2 - o.z) * idz;
    T rr_cap1 = (o.x + t_cap1 * d.x) * (o.x + t_cap1 * d.x) + (o.y + t_cap1 * d.y) * (o.y + t_cap1 * d.y);
    T rr_cap2 = (o.x + t_cap2 * d.x) * (o.x + t_cap2 * d.x) + (o.y + t_cap2 * d.y) * (o.y + t_cap2 * d.y);
    t_cap1 = rr_cap1 < r1r1 && t_cap1 > t_min ? t_cap1 : RT_MAX;
    t_cap2 = rr_cap2 < r2r2 && t_cap2 > t_min ? t_cap2 : RT_MAX;
    T c2 = d.x * d.x + d.y * d.y - d.z * d.z * tth2;
    T c1 = o.x * d.x + o.y * d.y - (o.z - z0) * d.z * tth2;
    T c0 = o.x * o.x + o.y * o.y - (o.z - z0) * (o.z - z0) * tth2;
    T t_near, t_far, disc, sdisc;
    if (fp64) {
        double tn, tf, di, sd;
        robust_quadratic_roots_disqualifying<double>(double(RT_MAX), tn, tf, di, sd, double(c2), double(c1), double(c0));
        t_near = T(tn); t_far = T(tf); disc = T(di); sdisc = T(sd);
    } else {
        robust_quadratic_roots_disqualifying(RT_MAX, t_near, t_far, disc, sdisc, c2, c1, c0);
    }
    T z_near = o.z + t_near * d.z;
    T z_far = o.z + t_far * d.z;
    t_near = z_near > z1 && z_near < z2 && t_near > t_min ? t_near : RT_MAX;
    t_far = z_far > z1 && z_far < z2 && t_far > t_min ? t_far : RT_MAX;
    T t_cand = fmn(fmn(t_near, t_far), fmn(t_cap1, t_cap2));
    valid = t_cand > t_min && t_cand < RT_MAX;
    if (valid) {
        if (t_cand == t_cap1 || t_cand == t_cap2) {
            is.x = T(0); is.y = T(0);
            is.z = t_cand == t_cap2 ? T(1) : T(-1);
        } else {
            V3<T> nn = normalize(v3<T>(o.x + t_cand * d.x, o.y + t_cand * d.y, (z0 - (o.z + t_cand * d.z)) * tth2));
            is.x = nn.x; is.y = nn.y; is.z = nn.z;
        }
        is.w = t_cand;
    }
}

// csg_intersect_leaf_convexpolyhedron.h:22-106
template <typename T>
void leaf_convexpolyhedron(bool& valid, Isect<T>& is, const Node& n, const float* plan, T t_min, V3<T> o, V3<T> d) {
    const T INF = std::numeric_limits<T>::infinity();
    T t0 = -INF, t1 = INF;
    V3<T> t0n = v3<T>(0, 0, 0), t1n = v3<T>(0, 0, 0);
    unsigned planeIdx = n.u(0), planeNum = n.u(1);
    for (unsigned i = 0; i < planeNum; i++) {
        const float* pl = plan + 4 * size_t(planeIdx + i);
        V3<T> nn = v3<T>(pl[0], pl[1], pl[2]);
        T dplane = pl[3];
        T nd = dot(nn, d);
        T no = dot(nn, o);
        T dist = no - dplane;
        T t_cand = -dist / nd;
        bool parallel_inside = nd == T(0) && dist < T(0);
        bool parallel_outside = nd == T(0) && dist > T(0);
        if (parallel_inside) continue;
        if (parallel_outside) { valid = false; return; }
        if (nd < T(0)) { if (t_cand > t0) { t0 = t_cand; t0n = nn; } }
        else { if (t_cand < t1) { t1 = t_cand; t1n = nn; } }
    }
    valid = t0 < t1;
    if (valid) {
        if (t0 > t_min) { is.x = t0n.x; is.y = t0n.y; is.z = t0n.z; is.w = t0; }
        else if (t1 > t_min) { is.x = t1n.x; is.y = t1n.y; is.z = t1n.z; is.w = t1; }
    }
}

// csg_intersect_leaf_phicut.h:581-635
template <typename T>
void leaf_phicut(bool& valid, Isect<T>& is, const Node& n, T t_min, V3<T> o, V3<T> d) {
    T cosPhi0 = n.f[0], sinPhi0 = n.f[1], cosPhi1 = n.f[2], sinPhi1 = n.f[3];
    T d_n0 = d.x * sinPhi0 + d.y * (-cosPhi0);
    T d_n1 = d.x * (-sinPhi1) + d.y * (cosPhi1);
    T o_n0 = o.x * sinPhi0 + o.y * (-cosPhi0);
    T o_n1 = o.x * (-sinPhi1) + o.y * (cosPhi1);
    T t0 = d_n0 == T(0) ? t_min : -o_n0 / d_n0;
    T t1 = d_n1 == T(0) ? t_min : -o_n1 / d_n1;
    T PR = d_n0 == T(0) ? -o_n0 : -d_n0;
    T QR = d_n1 == T(0) ? -o_n1 : -d_n1;
    T PQ = cosPhi0 * sinPhi1 - cosPhi1 * sinPhi0;
    bool unbounded_exit = PQ >= T(0) ? (PR >= T(0) && QR <= T(0)) : (PR >= T(0) || QR <= T(0));
    T side0 = o.x * cosPhi0 + o.y * sinPhi0 + (d.x * cosPhi0 + d.y * sinPhi0) * t0;
    T side1 = o.x * cosPhi1 + o.y * sinPhi1 + (d.x * cosPhi1 + d.y * sinPhi1) * t1;
    if (side0 < T(0)) t0 = t_min;
    if (side1 < T(0)) t1 = t_min;
    T t_near = fmn(t0, t1);
    T t_far = fmx(t0, t1);
    T t_cand = t_near > t_min ? t_near : (t_far > t_min ? t_far : t_min);
    valid = t_cand > t_min;
    if (valid) {
        is.x = t_cand == t1 ? -sinPhi1 : sinPhi0;
        is.y = t_cand == t1 ? cosPhi1 : -cosPhi0;
        is.z = T(0);
        is.w = t_cand;
    } else if (unbounded_exit) {
        is.y = -is.y;
    }
}

// csg_intersect_leaf_hyperboloid.h:18-106 : x^2+y^2 = r0^2((z/zf)^2+1) between z1 < z2
template <typename T>
void leaf_hyperboloid(bool& valid, Isect<T>& is, const Node& n, T t_min, V3<T> o, V3<T> dir) {
    const T zero(0), one(1), tmax(1.e27f);
    const T r0 = n.f[0], zf = n.f[1], z1 = n.f[2], z2 = n.f[3];
    const T rr0 = r0 * r0;
    const T z1s = z1 / zf, z2s = z2 / zf;
    const T rr1 = rr0 * (z1s * z1s + one), rr2 = rr0 * (z2s * z2s + one);
    const T A = -rr0 / (zf * zf);
    const T B = -rr0;
    const T sx = dir.x, sy = dir.y, sz = dir.z, ox = o.x, oy = o.y, oz = o.z;
    const T d = sx * sx + sy * sy + A * sz * sz;
    const T b = ox * sx + oy * sy + A * oz * sz;
    const T c = ox * ox + oy * oy + A * oz * oz + B;
    T t1hyp, t2hyp, disc, sdisc;
    robust_quadratic_roots(t1hyp, t2hyp, disc, sdisc, d, b, c);
    const T h1z = oz + t1hyp * sz, h2z = oz + t2hyp * sz;
    const T osz = one / sz;
    const T t2cap = (z2 - oz) * osz, t1cap = (z1 - oz) * osz;
    const V3<T> c1 = o + t1cap * dir, c2 = o + t2cap * dir;
    const T crr1 = c1.x * c1.x + c1.y * c1.y, crr2 = c2.x * c2.x + c2.y * c2.y;
    const T tc0 = t1hyp > t_min && disc > zero && h1z > z1 && h1z < z2 ? t1hyp : tmax;
    const T tc1 = t2hyp > t_min && disc > zero && h2z > z1 && h2z < z2 ? t2hyp : tmax;
    const T tc2 = t2cap > t_min && crr2 < rr2 ? t2cap : tmax;
    const T tc3 = t1cap > t_min && crr1 < rr1 ? t1cap : tmax;
    const T t_cand = std::fmin(std::fmin(tc0, tc1), std::fmin(tc2, tc3));
    valid = t_cand > t_min && t_cand < tmax;
    if (valid) {
        is.w = t_cand;
        if (t_cand == t1hyp || t_cand == t2hyp) {
            const V3<T> p = o + t_cand * dir;
            const V3<T> nn = normalize(v3<T>(p.x, p.y, A * p.z));
            is.x = nn.x; is.y = nn.y; is.z = nn.z;
        } else {
            is.x = zero; is.y = zero; is.z = t_cand == t1cap ? -one : one;
        }
    }
}

// csg_intersect_leaf_halfspace.h:155-187: halfspace p.n < w behind the plane p.n = w, q0 =
// (unit outward normal, w). t = -(o.n - w)/(d.n), or t_min for a ray parallel to the plane;
// a miss from inside (o.n - w < -1e-9) signals the exit at infinity with isect.y = -0.
template <typename T>
void leaf_halfspace(bool& valid, Isect<T>& is, const Node& n, T t_min, V3<T> o, V3<T> d) {
    V3<T> nn = v3<T>(n.f[0], n.f[1], n.f[2]);
    T w = n.f[3];
    T on = dot(o, nn), dn = dot(d, nn);
    T on_w = on - w;
    bool inside = on_w < T(-1e-9f);
    T t = std::fabs(dn) > T(0) ? -on_w / dn : t_min;
    valid = t > t_min;
    if (valid) { is.x = nn.x; is.y = nn.y; is.z = nn.z; is.w = t; }
    else if (inside) { is.y = -T(0); }
}

// qat4::right_multiply (sysrap/sqat4.h:45-51) for an itra/instance matrix stored as float[16]
template <typename T>
inline V3<T> right_multiply(const float* q, V3<T> v, T w) {
    return v3<T>(T(q[0]) * v.x + T(q[4]) * v.y + T(q[8]) * v.z + T(q[12]) * w,
                 T(q[1]) * v.x + T(q[5]) * v.y + T(q[9]) * v.z + T(q[13]) * w,
                 T(q[2]) * v.x + T(q[6]) * v.y + T(q[10]) * v.z + T(q[14]) * w);
}
// qat4::left_multiply_inplace(v, 0) (sysrap/sqat4.h:96-106)
template <typename T>
inline void left_multiply_normal(const float* q, Isect<T>& v) {
    T x = T(q[0]) * v.x + T(q[1]) * v.y + T(q[2]) * v.z;
    T y = T(q[4]) * v.x + T(q[5]) * v.y + T(q[6]) * v.z;
    T z = T(q[8]) * v.x + T(q[9]) * v.y + T(q[10]) * v.z;
    v.x = x; v.y = y; v.z = z;
}

// intersect_leaf (CSG/csg_intersect_leaf.h:175-324)
template <typename T>
void intersect_leaf(const Ctx& c, bool& valid, Isect<T>& is, const Node& n, T t_min, V3<T> ro, V3<T> rd) {
    valid = false;
    is = Isect<T>{0, 0, 0, 0};
    unsigned typecode = n.u(14);
    unsigned gt = n.u(15) & 0x7fffffffu;
    bool complement = (n.u(15) & 0x80000000u) != 0;
    const float* q = gt > 0 ? c.itra + 16 * size_t(gt - 1) : nullptr;
    V3<T> o = q ? right_multiply(q, ro, T(1)) : ro;
    V3<T> d = q ? right_multiply(q, rd, T(0)) : rd;
    bool fp64 = c.cfg.fp64_roots != 0;
    switch (typecode) {
        case CSG_SPHERE: leaf_sphere(valid, is, n, t_min, o, d, fp64); break;
        case CSG_ZSPHERE: leaf_zsphere(valid, is, n, t_min, o, d, fp64); break;
        case CSG_CYLINDER: leaf_cylinder(valid, is, n, t_min, o, d, fp64); break;
        case CSG_BOX3: leaf_box3(valid, is, n, t_min, o, d); break;
        case CSG_CONE: leaf_cone(valid, is, n, t_min, o, d, fp64); break;
        case CSG_CONVEXPOLYHEDRON: leaf_convexpolyhedron(valid, is, n, c.plan, t_min, o, d); break;
        case CSG_PHICUT: leaf_phicut(valid, is, n, t_min, o, d); break;
        case CSG_HALFSPACE: leaf_halfspace(valid, is, n, t_min, o, d); break;
        case CSG_HYPERBOLOID: leaf_hyperboloid(valid, is, n, t_min, o, d); break;
        default: break;
    }
    if (valid && q) left_multiply_normal(q, is);
    if (complement) {
        is.x = valid ? -is.x : -T(0);
        is.y = valid ? -is.y : is.y;
        is.z = valid ? -is.z : is.z;
    }
}

// ---------------------------------------------------------------- CSG tree
// csg_classify.h:55 CSG_CLASSIFY, :86-101 LUT
enum { State_Enter = 0, State_Exit = 1, State_Miss = 2 };
enum { CTRL_RETURN_MISS = 0, CTRL_RETURN_A = 1, CTRL_RETURN_B = 2, CTRL_RETURN_FLIP_B = 3,
       CTRL_LOOP_A = 4, CTRL_LOOP_B = 5 };

template <typename T>
inline int classify(const Isect<T>& ise, V3<T> dir, T tmin) {
    return std::fabs(ise.w) > tmin ? ((ise.x * dir.x + ise.y * dir.y + ise.z * dir.z) < T(0) ? State_Enter : State_Exit)
                                   : State_Miss;
}

// The boolean decision table of Kensler's single-hit CSG as packed in csg_classify.h:89-90,
// restated as explicit tables [op][stateA][stateB] for A-closer and B-closer.
inline int lut(int op, int a, int b, bool acloser) {
    static const uint32_t AC[3] = {0x22121141u, 0x00014014u, 0x00141141u};
    static const uint32_t BC[3] = {0x22115122u, 0x00022055u, 0x00133155u};
    unsigned offset = 3u * unsigned(a) + unsigned(b);
    unsigned index = unsigned(op) - 1u;
    const uint32_t* t = acloser ? AC : BC;
    return offset < 8 ? int((t[index] >> (offset * 4)) & 0xf) : CTRL_RETURN_MISS;
}

inline unsigned tree_height(unsigned numNode) { return unsigned(__builtin_ffs(int(numNode + 1)) - 2); }
inline unsigned tree_depth(unsigned idx) { return 31u - unsigned(__builtin_clz(idx)); }
inline unsigned postorder_next(unsigned i, unsigned elevation) {
    return (i & 1u) ? i >> 1 : (i << elevation) + (1u << elevation);
}
inline unsigned pack4(unsigned c, unsigned d) { return ((c & 0xffu) << 16) | ((d & 0xffu) << 24); }

// ---------------------------------------------------------------- list nodes
// subNum leaf nodes starting subOffset nodes after the prim's first node `root`
// (CSG/csg_intersect_node.h); `is` is written only for a valid intersect.
constexpr float RT_DEFAULT_MAX = 1.e27f;

// intersect_node_discontiguous (csg_intersect_node.h:655-686)
template <typename T>
void intersect_node_discontiguous(const Ctx& c, bool& valid, Isect<T>& is, const Node* nd, const Node* root, T t_min,
                                  V3<T> ro, V3<T> rd) {
    unsigned num_sub = nd->u(0), offset_sub = nd->u(1);
    Isect<T> closest{0, 0, 0, T(RT_DEFAULT_MAX)};
    Isect<T> sub{0, 0, 0, 0};                       // one per call: a leaf that reports valid
    for (unsigned isub = 0; isub < num_sub; isub++) { // without writing leaves the last values
        bool v = false;
        intersect_leaf(c, v, sub, root[offset_sub + isub], t_min, ro, rd);
        if (v && sub.w < closest.w) closest = sub;
    }
    valid = closest.w < T(RT_DEFAULT_MAX);
    if (valid) is = closest;
}

// intersect_node_contiguous (csg_intersect_node.h:393-624): zeroth pass for the nearest
// enter and the exit count; from inside, t-ordered enters extend the farthest exit
template <typename T>
void intersect_node_contiguous(const Ctx& c, bool& valid, Isect<T>& is, const Node* nd, const Node* root, T t_min,
                               V3<T> ro, V3<T> rd) {
    const int num_sub = int(nd->u(0)), offset_sub = int(nd->u(1));
    Isect<T> nearest_enter{0, 0, 0, T(RT_DEFAULT_MAX)};
    Isect<T> farthest_exit{0, 0, 0, t_min};
    int exit_count = 0;
    const T propagate_epsilon = T(0.0001f);
    Isect<T> sub{0, 0, 0, 0};                       // shared by all three passes
    for (int i = 0; i < num_sub; i++) {
        bool v = false;
        intersect_leaf(c, v, sub, root[offset_sub + i], t_min, ro, rd);
        if (v) {
            int st = classify(sub, rd, t_min);
            if (st == State_Enter) { if (sub.w < nearest_enter.w) nearest_enter = sub; }
            else if (st == State_Exit) exit_count += 1;
        }
    }
    if (exit_count == 0) {
        valid = nearest_enter.w > t_min && nearest_enter.w < T(RT_DEFAULT_MAX);
        if (valid) is = nearest_enter;
        return;
    }
    int enter_count = 0;
    T enter[8];
    int aux[8], idx[8];
    for (int isub = 0; isub < num_sub; isub++) {
        bool v = false;
        intersect_leaf(c, v, sub, root[offset_sub + isub], t_min, ro, rd);
        if (v) {
            int st = classify(sub, rd, t_min);
            if (st == State_Enter) {
                aux[enter_count] = isub; idx[enter_count] = enter_count; enter[enter_count] = sub.w;
                enter_count += 1;
            } else if (st == State_Exit) {
                exit_count += 1;
                if (sub.w > farthest_exit.w) farthest_exit = sub;
            }
        }
    }
    for (int i = 1; i < enter_count; i++) {          // insertion sort of indices by enter t
        int key = idx[i];
        int j = i - 1;
        while (j >= 0 && enter[idx[j]] > enter[key]) { idx[j + 1] = idx[j]; j = j - 1; }
        idx[j + 1] = key;
    }
    for (int i = 0; i < enter_count; i++) {
        T tminAdvanced = enter[idx[i]] + propagate_epsilon;
        int isub = aux[idx[i]];
        if (tminAdvanced < farthest_exit.w) {
            bool v = false;
            intersect_leaf(c, v, sub, root[offset_sub + isub], tminAdvanced, ro, rd);
            if (v && classify(sub, rd, tminAdvanced) == State_Exit) {
                exit_count += 1;
                if (sub.w > farthest_exit.w) farthest_exit = sub;
            }
        }
    }
    valid = exit_count > 0 && farthest_exit.w > t_min;
    if (valid) is = farthest_exit;
}

// intersect_node_overlap (csg_intersect_node.h:820-895)
template <typename T>
void intersect_node_overlap(const Ctx& c, bool& valid, Isect<T>& is, const Node* nd, const Node* root, T t_min,
                            V3<T> ro, V3<T> rd) {
    unsigned num_sub = nd->u(0), offset_sub = nd->u(1);
    Isect<T> farthest_enter{0, 0, 0, t_min};
    Isect<T> nearest_exit{0, 0, 0, T(RT_DEFAULT_MAX)};
    const T propagate_epsilon = T(0.0001f);
    unsigned enter_count = 0, exit_count = 0;
    Isect<T> sub{0, 0, 0, 0};                       // one per call, reused for the exit search
    for (unsigned isub = 0; isub < num_sub; isub++) {
        const Node& sn = root[offset_sub + isub];
        bool v = false;
        intersect_leaf(c, v, sub, sn, t_min, ro, rd);
        if (!v) continue;
        int st = classify(sub, rd, t_min);
        if (st == State_Enter) {
            enter_count += 1;
            if (sub.w > farthest_enter.w) farthest_enter = sub;
            T tminAdvanced = sub.w + propagate_epsilon;
            v = false;
            intersect_leaf(c, v, sub, sn, tminAdvanced, ro, rd);
            if (v && classify(sub, rd, tminAdvanced) == State_Exit) {
                exit_count += 1;
                if (sub.w < nearest_exit.w) nearest_exit = sub;
            }
        } else if (st == State_Exit) {
            exit_count += 1;
            if (sub.w < nearest_exit.w) nearest_exit = sub;
        }
    }
    valid = false;
    bool overlap_all = farthest_enter.w < nearest_exit.w && std::max(enter_count, exit_count) == num_sub;
    if (overlap_all) {
        if (farthest_enter.w > t_min && farthest_enter.w < T(RT_DEFAULT_MAX)) { valid = true; is = farthest_enter; }
        else if (nearest_exit.w > t_min && nearest_exit.w < T(RT_DEFAULT_MAX)) { valid = true; is = nearest_exit; }
    }
}

template <typename T>
void intersect_node(const Ctx& c, bool& valid, Isect<T>& is, const Node* nd, const Node* root, T tmin, V3<T> ro, V3<T> rd);

// intersect_tree (CSG/csg_intersect_tree.h:277-672)
template <typename T>
void intersect_tree(const Ctx& c, bool& valid, Isect<T>& is, const Node* root, T t_min, V3<T> ro, V3<T> rd) {
    const unsigned numNode = root->u(0);
    const unsigned height = tree_height(numNode);
    const T propagate_epsilon = T(0.0001f);
    int ierr = 0;
    struct { T tmin[4]; unsigned slice[4]; int curr; } tr;
    tr.curr = -1;
    struct { Isect<T> data[15]; int curr; } st;
    st.curr = -1;
    auto tpush = [&](unsigned s, T t) { if (tr.curr >= 3) return 1; tr.curr++; tr.slice[tr.curr] = s; tr.tmin[tr.curr] = t; return 0; };
    auto spush = [&](const Isect<T>& v) { if (st.curr >= 14) return 1; st.curr++; st.data[st.curr] = v; return 0; };
    auto spop = [&]() { if (st.curr < 0) return 1; st.curr--; return 0; };
    tpush(pack4(1u << height, 0), t_min);
    while (tr.curr > -1) {
        unsigned slice = tr.slice[tr.curr];
        T tmin = tr.tmin[tr.curr];
        tr.curr--;
        unsigned beginIdx = (slice >> 16) & 0xff, endIdx = (slice >> 24) & 0xff;
        unsigned nodeIdx = beginIdx;
        while (nodeIdx != endIdx) {
            unsigned depth = tree_depth(nodeIdx);
            unsigned elevation = height - depth;
            const Node* nd = root + nodeIdx - 1;
            unsigned typecode = nd->u(14);
            if (typecode == CSG_ZERO) { nodeIdx = postorder_next(nodeIdx, elevation); continue; }
            bool node_or_leaf = typecode >= CSG_NODE;
            if (node_or_leaf) {
                Isect<T> ni{0, 0, 0, 0};
                bool nv = false;
                intersect_node(c, nv, ni, nd, root, tmin, ro, rd);
                ni.w = std::copysign(ni.w, nodeIdx % 2 == 0 ? T(-1) : T(1));
                ierr = spush(ni);
                if (ierr) break;
            } else {
                if (st.curr < 1) { ierr = 1; break; }
                bool firstLeft = std::signbit(st.data[st.curr].w);
                bool secondLeft = std::signbit(st.data[st.curr - 1].w);
                if (!(firstLeft ^ secondLeft)) { ierr = 1; break; }
                int left = firstLeft ? st.curr : st.curr - 1;
                int right = firstLeft ? st.curr - 1 : st.curr;
                int l_state = classify(st.data[left], rd, tmin);
                int r_state = classify(st.data[right], rd, tmin);
                T t_left = std::fabs(st.data[left].w);
                T t_right = std::fabs(st.data[right].w);
                bool leftIsCloser = t_left <= t_right;
                bool l_complement = std::signbit(st.data[left].x);
                bool r_complement = std::signbit(st.data[right].x);
                bool l_unbounded = std::signbit(st.data[left].y);
                bool r_unbounded = std::signbit(st.data[right].y);
                bool l_promote = l_state == State_Miss && (l_complement || l_unbounded);
                bool r_promote = r_state == State_Miss && (r_complement || r_unbounded);
                if (r_promote) { r_state = State_Exit; leftIsCloser = true; }
                if (l_promote) { l_state = State_Exit; leftIsCloser = false; }
                int ctrl = lut(int(typecode), l_state, r_state, leftIsCloser);
                if (ctrl < CTRL_LOOP_A) {
                    Isect<T> result = ctrl == CTRL_RETURN_MISS ? Isect<T>{0, 0, 0, 0} : st.data[ctrl == CTRL_RETURN_A ? left : right];
                    if (ctrl == CTRL_RETURN_FLIP_B) { result.x = -result.x; result.y = -result.y; result.z = -result.z; }
                    result.w = std::copysign(result.w, nodeIdx % 2 == 0 ? T(-1) : T(1));
                    if ((ierr = spop())) break;
                    if ((ierr = spop())) break;
                    if ((ierr = spush(result))) break;
                } else {
                    int loopside = ctrl == CTRL_LOOP_A ? left : right;
                    int otherside = ctrl == CTRL_LOOP_A ? right : left;
                    unsigned leftIdx = 2 * nodeIdx, rightIdx = leftIdx + 1;
                    unsigned otherIdx = ctrl == CTRL_LOOP_A ? rightIdx : leftIdx;
                    (void)otherIdx;
                    T tminAdvanced = std::fabs(st.data[loopside].w) + propagate_epsilon;
                    Isect<T> other = st.data[otherside];
                    if ((ierr = spop())) break;
                    if ((ierr = spop())) break;
                    if ((ierr = spush(other))) break;
                    unsigned endTree = pack4(nodeIdx, endIdx);
                    unsigned leftTree = pack4(leftIdx << (elevation - 1), rightIdx << (elevation - 1));
                    unsigned rightTree = pack4(rightIdx << (elevation - 1), nodeIdx);
                    unsigned loopTree = ctrl == CTRL_LOOP_A ? leftTree : rightTree;
                    if ((ierr = tpush(endTree, tmin))) break;
                    if ((ierr = tpush(loopTree, tminAdvanced))) break;
                    break;
                }
            }
            nodeIdx = postorder_next(nodeIdx, elevation);
        }
        if (ierr) break;
    }
    if (st.curr == 0) is = st.data[0];
    valid = is.w > T(0);
}

// intersect_node_discontiguous / overlap / contiguous are list types not used by the
// benchmark geometry; intersect_node dispatches leaves (csg_intersect_node.h:914-938).
template <typename T>
void intersect_node(const Ctx& c, bool& valid, Isect<T>& is, const Node* nd, const Node* root, T tmin, V3<T> ro, V3<T> rd) {
    valid = false;
    unsigned tc = nd->u(14);
    if (tc == CSG_CONTIGUOUS) intersect_node_contiguous(c, valid, is, nd, root, tmin, ro, rd);
    else if (tc == CSG_DISCONTIGUOUS) intersect_node_discontiguous(c, valid, is, nd, root, tmin, ro, rd);
    else if (tc == CSG_OVERLAP) intersect_node_overlap(c, valid, is, nd, root, tmin, ro, rd);
    else intersect_leaf(c, valid, is, *nd, tmin, ro, rd);
}

// intersect_prim (CSG/csg_intersect_tree.h:684-719)
template <typename T>
bool intersect_prim(const Ctx& c, Isect<T>& is, const Node* root, T t_min, V3<T> ro, V3<T> rd) {
    unsigned typecode = root->u(14);
    bool valid = false;
    is = Isect<T>{0, 0, 0, 0};
    if (typecode >= CSG_LEAF) intersect_leaf(c, valid, is, *root, t_min, ro, rd);
    else if (typecode < CSG_NODE) intersect_tree(c, valid, is, root, t_min, ro, rd);
    else intersect_node(c, valid, is, root, root, t_min, ro, rd);
    return valid;
}

// ---------------------------------------------------------------- instances
// Inverse of an instance transform (identity info in the .w column is ignored), double.
inline Q4x3<double> invert_instance(const float* q) {
    double a[3][3], t[3];
    for (int r = 0; r < 3; r++) for (int k = 0; k < 3; k++) a[r][k] = q[4 * r + k];
    for (int k = 0; k < 3; k++) t[k] = q[12 + k];
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    double inv[3][3];
    inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
    inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
    inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
    inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
    Q4x3<double> r;
    // row convention: p_o = (p_w - t) * inv
    for (int i = 0; i < 3; i++) for (int k = 0; k < 3; k++) r.m[i][k] = inv[i][k];
    for (int k = 0; k < 3; k++) r.m[3][k] = -(t[0] * inv[0][k] + t[1] * inv[1][k] + t[2] * inv[2][k]);
    return r;
}

template <typename T>
struct PRD {
    V3<T> normal; T t;
    float lposcost, lposfphi;
    uint32_t iindex_identity, gp_boundary;
};

template <typename T> inline T mat(const Q4x3<double>& m, int r, int k) { return T(m.m[r][k]); }
template <> inline float mat<float>(const Q4x3<double>& m, int r, int k) { return float(m.m[r][k]); }

// Closest hit over every instance and every prim of its solid: the optixTrace contract
// of CSGOptiX7.cu:110-216 (closest hit in (tmin,tmax], anyhit disabled) with the IS/CH/MS
// programs of CSGOptiX7.cu:655-940. Ties at equal t resolve to the smaller
// (instance index, prim index), the engine's deterministic rule.
// conservative slab test in double against a box expanded by 1e-4 relative + 1e-2 mm
inline bool box_may_hit(const double* b, double ox, double oy, double oz, double dx, double dy, double dz, double tmax) {
    double t0 = 0.0, t1 = tmax;
    const double o[3] = {ox, oy, oz}, d[3] = {dx, dy, dz};
    for (int k = 0; k < 3; k++) {
        double m = std::max(std::fabs(b[k]), std::fabs(b[3 + k]));
        double lo = b[k] - 1e-4 * m - 1e-2, hi = b[3 + k] + 1e-4 * m + 1e-2;
        if (d[k] == 0.0) { if (o[k] < lo || o[k] > hi) return false; continue; }
        double a = (lo - o[k]) / d[k], bb = (hi - o[k]) / d[k];
        if (a > bb) std::swap(a, bb);
        t0 = std::max(t0, a); t1 = std::min(t1, bb);
        if (t0 > t1) return false;
    }
    return true;
}

template <typename T>
bool trace(const Ctx& c, V3<T> pos, V3<T> mom, T tmin, T tmax, PRD<T>& prd) {
    bool found = false;
    T best = tmax;
    uint64_t best_key = ~0ull;
    V3<T> best_n{0, 0, 0};
    V3<T> best_lpos{0, 0, 0};
    uint32_t best_ii = 0, best_gpbd = 0;
    for (int64_t i = 0; i < c.num_inst; i++) {
        const float* q = c.inst + 16 * i;
        int32_t gas = 0; std::memcpy(&gas, q + 7, 4);
        int32_t sid1 = 0; std::memcpy(&sid1, q + 11, 4);
        const int32_t* so = c.solid + 12 * size_t(gas);
        int numPrim = so[4], primOffset = so[5];
        // instance visibility mask, one bit per solid up to the eighth (Properties.cc:45-55,
        // SBT.cc:538-545), against the ray's mask: OptiX skips an instance sharing no bit
        if (((1u << std::min<int32_t>(gas, 7)) & c.cfg.vizmask) == 0u) continue;
        bool ident = c.inst_identity_transform[i] != 0;
        if (c.cull && !box_may_hit(&c.inst_box[6 * size_t(i)], double(pos.x), double(pos.y), double(pos.z),
                                   double(mom.x), double(mom.y), double(mom.z), double(best))) continue;
        V3<T> o = pos, d = mom;
        const Q4x3<double>& w = c.w2o[i];
        if (!ident) {
            o = v3<T>(mat<T>(w, 0, 0) * pos.x + mat<T>(w, 1, 0) * pos.y + mat<T>(w, 2, 0) * pos.z + mat<T>(w, 3, 0),
                      mat<T>(w, 0, 1) * pos.x + mat<T>(w, 1, 1) * pos.y + mat<T>(w, 2, 1) * pos.z + mat<T>(w, 3, 1),
                      mat<T>(w, 0, 2) * pos.x + mat<T>(w, 1, 2) * pos.y + mat<T>(w, 2, 2) * pos.z + mat<T>(w, 3, 2));
            d = v3<T>(mat<T>(w, 0, 0) * mom.x + mat<T>(w, 1, 0) * mom.y + mat<T>(w, 2, 0) * mom.z,
                      mat<T>(w, 0, 1) * mom.x + mat<T>(w, 1, 1) * mom.y + mat<T>(w, 2, 1) * mom.z,
                      mat<T>(w, 0, 2) * mom.x + mat<T>(w, 1, 2) * mom.y + mat<T>(w, 2, 2) * mom.z);
        }
        for (int p = 0; p < numPrim; p++) {
            const float* pr = c.prim + 16 * size_t(primOffset + p);
            int32_t nodeOffset; std::memcpy(&nodeOffset, pr + 1, 4);
            uint32_t gpi = f2u(pr[15]);
            const Node* root = c.node + nodeOffset;
            if (c.cull && !box_may_hit(&c.prim_box[6 * size_t(primOffset + p)], double(o.x), double(o.y), double(o.z),
                                       double(d.x), double(d.y), double(d.z), double(best))) continue;
            Isect<T> is;
            if (!intersect_prim(c, is, root, tmin, o, d)) continue;
            T t = is.w;
            // the hit is reported with optixReportIntersection (CSGOptiX/CSGOptiX7.cu:912),
            // which ignores a distance outside the ray interval: a convex polyhedron wholly
            // behind the ray reports valid with the zeroed isect, t = 0 < tmin
            if (!(t >= tmin)) continue;
            uint64_t key = (uint64_t(i) << 32) | uint64_t(p);
            if (t < best || (found && t == best && key < best_key)) {
                found = true;
                best = t; best_key = key;
                best_n = v3<T>(is.x, is.y, is.z);
                best_lpos = v3<T>(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
                uint32_t bnd = root->u(6);
                best_gpbd = ((gpi & 0xffffu) << 16) | (bnd & 0xffffu);
                best_ii = ((uint32_t(i) & 0xffffu) << 16) | (uint32_t(sid1) & 0xffffu);
                if (!ident) {
                    T nx = mat<T>(w, 0, 0) * is.x + mat<T>(w, 0, 1) * is.y + mat<T>(w, 0, 2) * is.z;
                    T ny = mat<T>(w, 1, 0) * is.x + mat<T>(w, 1, 1) * is.y + mat<T>(w, 1, 2) * is.z;
                    T nz = mat<T>(w, 2, 0) * is.x + mat<T>(w, 2, 1) * is.y + mat<T>(w, 2, 2) * is.z;
                    best_n = v3<T>(nx, ny, nz);
                }
            }
        }
    }
    if (!found) {
        prd.normal = v3<T>(0, 0, 0); prd.t = T(0);
        prd.lposcost = 0.f; prd.lposfphi = 0.f;
        prd.iindex_identity = 0xffffffffu; prd.gp_boundary = 0xffffffffu;
        return false;
    }
    prd.normal = best_n; prd.t = best;
    T ll = std::sqrt(dot(best_lpos, best_lpos));
    prd.lposcost = float(best_lpos.z / ll);
    prd.lposfphi = float((xatan2(best_lpos.y, best_lpos.x) + pi<T>()) / (T(2) * pi<T>()));
    prd.iindex_identity = best_ii;
    prd.gp_boundary = best_gpbd;
    return true;
}

// trace<refine> (CSGOptiX/CSGOptiX7.cu:110-216) with the miss program's prd distance 1
// (CSGOptiX7.cu:655-682, WITH_PRD): with PropagateRefine, when t_approx = 0.99 * distance
// exceeds PropagateRefineDistance the ray is traced again from pos + t_approx * mom and
// t_approx is added to the second distance (quad2::distance_add, sysrap/squad.h:235).
template <typename T>
bool trace_refined(const Ctx& c, V3<T> pos, V3<T> mom, T tmin, T tmax, PRD<T>& prd) {
    bool found = trace<T>(c, pos, mom, tmin, tmax, prd);
    if (!found) prd.t = T(1);
    if (!c.cfg.propagate_refine) return found;
    const T t_approx = T(0.99) * prd.t;
    if (!(t_approx > T(c.cfg.refine_distance))) return found;
    const V3<T> closer = v3<T>(pos.x + t_approx * mom.x, pos.y + t_approx * mom.y, pos.z + t_approx * mom.z);
    found = trace<T>(c, closer, mom, tmin, tmax, prd);
    if (!found) prd.t = T(1);
    prd.t = prd.t + t_approx;
    return found;
}

// ---------------------------------------------------------------- photon
template <typename T>
struct Photon {
    V3<T> pos; T time;
    V3<T> mom; uint32_t hitcount_iindex;
    V3<T> pol; T wavelength;
    uint32_t orient_boundary_flag, identity, index, flagmask;
    unsigned flag() const { return orient_boundary_flag & 0xffffu; }
    unsigned boundary() const { return (orient_boundary_flag & 0x7fff0000u) >> 16; }
    void set_flag(unsigned f) { orient_boundary_flag = (orient_boundary_flag & 0xffff0000u) | (f & 0xffffu); flagmask |= f; }
    void zero_flags() { orient_boundary_flag = 0; identity = 0; index = 0; flagmask = 0; hitcount_iindex = 0; }
    // sphoton::set_prd (sysrap/sphoton.h:482-488)
    void set_prd(unsigned boundary, unsigned id, T orient, unsigned iindex) {
        orient_boundary_flag = (orient_boundary_flag & 0x8000ffffu) | ((boundary & 0x7fffu) << 16);
        identity = (identity & 0xff000000u) | (id & 0x00ffffffu);
        orient_boundary_flag = (orient_boundary_flag & 0x7fffffffu) | ((orient < T(0) ? 1u : 0u) << 31);
        hitcount_iindex = 0x00010000u | (iindex & 0xffffu);
    }
    void set_index(uint64_t full) {
        index = uint32_t(full & 0xffffffffu);
        identity = (uint32_t((full >> 32) & 0xffu) << 24) | (identity & 0xffffffu);
    }
    void store(float* o) const {
        o[0] = float(pos.x); o[1] = float(pos.y); o[2] = float(pos.z); o[3] = float(time);
        o[4] = float(mom.x); o[5] = float(mom.y); o[6] = float(mom.z); o[7] = u2f(hitcount_iindex);
        o[8] = float(pol.x); o[9] = float(pol.y); o[10] = float(pol.z); o[11] = float(wavelength);
        o[12] = u2f(orient_boundary_flag); o[13] = u2f(identity); o[14] = u2f(index); o[15] = u2f(flagmask);
    }
    void load(const float* o) {
        pos = v3<T>(o[0], o[1], o[2]); time = o[3];
        mom = v3<T>(o[4], o[5], o[6]); hitcount_iindex = f2u(o[7]);
        pol = v3<T>(o[8], o[9], o[10]); wavelength = o[11];
        orient_boundary_flag = f2u(o[12]); identity = f2u(o[13]); index = f2u(o[14]); flagmask = f2u(o[15]);
    }
};

template <typename T>
struct State { V4<T> material1, m1group2, material2, surface; uint32_t optical[4]; };

// smath::rotateUz (sysrap/smath.h:77-96)
template <typename T>
inline void rotateUz(V3<T>& d, V3<T> u) {
    T up = u.x * u.x + u.y * u.y;
    if (up > T(0)) {
        up = std::sqrt(up);
        T px = d.x, py = d.y, pz = d.z;
        d.x = (u.x * u.z * px - u.y * py) / up + u.x * pz;
        d.y = (u.y * u.z * px + u.x * py) / up + u.y * pz;
        d.z = -up * px + u.z * pz;
    } else if (u.z < T(0)) {
        d.x = -d.x; d.z = -d.z;
    }
}

// rejection loops stop after this many turns (the engine applies the same cap)
constexpr int MAX_REJECT = 10000;

template <typename T>
struct Sim {
    const Ctx& c;
    Stream rng;
    T U() { return T(rng.next()); }
    explicit Sim(const Ctx& c_) : c(c_) {}
    bool aligned() const { return c.cfg.rng_schedule == OXR_RNG_ALIGNED; }

    // qbnd::fill_state (qudarap/qbnd.h:184-214)
    void fill_state(State<T>& s, unsigned boundary, T wavelength, T cosTheta) {
        int line = int(boundary) * 4;
        int m1_line = cosTheta > T(0) ? line + 3 : line + 0;
        int m2_line = cosTheta > T(0) ? line + 0 : line + 3;
        int su_line = cosTheta > T(0) ? line + 2 : line + 1;
        s.material1 = boundary_lookup<T>(c, wavelength, m1_line, 0);
        s.m1group2 = boundary_lookup<T>(c, wavelength, m1_line, 1);
        s.material2 = boundary_lookup<T>(c, wavelength, m2_line, 0);
        s.surface = boundary_lookup<T>(c, wavelength, su_line, 0);
        for (int k = 0; k < 4; k++) s.optical[k] = c.optical[4 * su_line + k];
    }

    // qsim::uniform_sphere (qudarap/qsim.h:177-183)
    V3<T> uniform_sphere(T u0, T u1) {
        T phi = u0 * T(2) * pi<T>();
        T cosTheta = T(2) * u1 - T(1);
        T sinTheta = std::sqrt(T(1) - cosTheta * cosTheta);
        return v3<T>(xcos(phi) * sinTheta, xsin(phi) * sinTheta, cosTheta);
    }

    // qsim::rayleigh_scatter (qudarap/qsim.h:601-689)
    void rayleigh_scatter(Photon<T>& p) {
        V3<T> direction, polarization;
        bool looping = true;
        int turns = 0;
        do {
            T u0 = U(), u1 = U(), u2 = U(), u3 = U(), u4 = U();
            T cosTheta = u0;
            T sinTheta = std::sqrt(T(1) - u0 * u0);
            if (u1 < T(0.5)) cosTheta = -cosTheta;
            T ph = T(2) * pi<T>() * u2;
            T sinPhi = xsin(ph), cosPhi = xcos(ph);
            direction = v3<T>(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
            rotateUz(direction, p.mom);
            T constant = -dot(direction, p.pol);
            polarization = v3<T>(p.pol.x + constant * direction.x, p.pol.y + constant * direction.y,
                                 p.pol.z + constant * direction.z);
            if (dot(polarization, polarization) == T(0)) {
                T ph3 = T(2) * pi<T>() * u3;
                sinPhi = xsin(ph3); cosPhi = xcos(ph3);
                polarization = v3<T>(cosPhi, sinPhi, T(0));
                rotateUz(polarization, direction);
            } else {
                if (u3 < T(0.5)) polarization = -polarization;
            }
            polarization = normalize(polarization);
            T doCosTheta = dot(polarization, p.pol);
            T doCosTheta2 = doCosTheta * doCosTheta;
            looping = doCosTheta2 < u4;
        } while (looping && ++turns < MAX_REJECT);
        p.mom = direction;
        p.pol = polarization;
    }

    // qsim::propagate_to_boundary (qudarap/qsim.h:718-863)
    int propagate_to_boundary(unsigned& flag, Photon<T>& p, const State<T>& s, T distance_to_boundary) {
        T absorption_length = s.material1.y;
        T scattering_length = s.material1.z;
        T reemission_prob = s.material1.w;
        T group_velocity = s.m1group2.x;
        if (aligned()) { U(); U(); }
        T u_scattering = U();
        T u_absorption = U();
        T scattering_distance = -scattering_length * xlog(u_scattering);
        T absorption_distance = -absorption_length * xlog(u_absorption);
        if (absorption_distance <= scattering_distance) {
            if (absorption_distance <= distance_to_boundary) {
                p.time += absorption_distance / group_velocity;
                p.pos = p.pos + absorption_distance * p.mom;
                T u_reemit = reemission_prob == T(0) ? T(2) : U();
                if (u_reemit < reemission_prob) {
                    T u_wl = U(), u_mph = U(), u_mct = U(), u_pph = U(), u_pct = U();
                    p.wavelength = scint_wavelength<T>(c, u_wl);
                    p.mom = uniform_sphere(u_mph, u_mct);
                    p.pol = normalize(cross(uniform_sphere(u_pph, u_pct), p.mom));
                    flag = BULK_REEMIT;
                    return CONTINUE;
                } else {
                    flag = BULK_ABSORB;
                    return BREAK;
                }
            }
        } else {
            if (scattering_distance <= distance_to_boundary) {
                p.time += scattering_distance / group_velocity;
                p.pos = p.pos + scattering_distance * p.mom;
                rayleigh_scatter(p);
                flag = BULK_SCATTER;
                return CONTINUE;
            }
        }
        p.pos = p.pos + distance_to_boundary * p.mom;
        p.time += distance_to_boundary / group_velocity;
        return BOUNDARY;
    }

    // qsim::propagate_at_boundary (qudarap/qsim.h:998-1205), theTransmittance = -1
    int propagate_at_boundary(unsigned& flag, Photon<T>& p, const State<T>& s, V3<T> normal) {
        T n1 = s.material1.x;
        T n2 = s.material2.x;
        T eta = n1 / n2;
        T _c1 = -dot(p.mom, normal);
        V3<T> oriented_normal = _c1 < T(0) ? -normal : normal;
        V3<T> trans = cross(p.mom, oriented_normal);
        T trans_length = length(trans);
        bool normal_incidence = trans_length < T(1e-6f);
        V3<T> A_trans = normal_incidence ? p.pol : trans / trans_length;
        T E1_perp = dot(p.pol, A_trans);
        T c1 = std::fabs(_c1);
        T c2c2 = T(1) - eta * eta * (T(1) - c1 * c1);
        bool tir = c2c2 < T(0);
        T EdotN = dot(p.pol, oriented_normal);
        T c2 = tir ? T(0) : std::sqrt(c2c2);
        T n1c1 = n1 * c1, n2c2 = n2 * c2, n2c1 = n2 * c1, n1c2 = n1 * c2;
        T E1x = normal_incidence ? T(0) : E1_perp;
        T E1y = normal_incidence ? T(1) : length(p.pol - (E1_perp * A_trans));
        T E2tx = T(2) * n1c1 * E1x / (n1c1 + n2c2);
        T E2ty = T(2) * n1c1 * E1y / (n2c1 + n1c2);
        T E2rx = E2tx - E1x;
        T E2ry = (n2 * E2ty / n1) - E1y;
        T rinv = T(1) / std::sqrt(E2rx * E2rx + E2ry * E2ry);
        T RRx = E2rx * rinv, RRy = E2ry * rinv;
        T tinv = T(1) / std::sqrt(E2tx * E2tx + E2ty * E2ty);
        T TTx = E2tx * tinv, TTy = E2ty * tinv;
        T TransCoeff = (tir || n1c1 == T(0)) ? T(0) : n2c2 * (E2tx * E2tx + E2ty * E2ty) / n1c1;
        if (aligned()) U();
        T u_reflect = U();
        bool reflect = u_reflect > TransCoeff;
        p.mom = reflect ? p.mom + T(2) * c1 * oriented_normal
                        : eta * p.mom + (eta * c1 - c2) * oriented_normal;
        V3<T> A_paral = normalize(cross(p.mom, A_trans));
        if (normal_incidence) {
            p.pol = reflect ? p.pol * (n2 > n1 ? T(-1) : T(1)) : p.pol;
        } else if (reflect) {
            p.pol = tir ? -p.pol + T(2) * EdotN * oriented_normal : RRx * A_trans + RRy * A_paral;
        } else {
            p.pol = TTx * A_trans + TTy * A_paral;
        }
        flag = reflect ? BOUNDARY_REFLECT : BOUNDARY_TRANSMIT;
        if (aligned() && flag == BOUNDARY_REFLECT) { U(); U(); U(); U(); }
        return CONTINUE;
    }

    // qsim::random_direction_marsaglia (qudarap/qsim.h:551-572)
    V3<T> random_direction_marsaglia() {
        T u, v, b, a;
        int turns = 0;
        do {
            T u0 = U(), u1 = U();
            u = T(2) * u0 - T(1);
            v = T(2) * u1 - T(1);
            b = u * u + v * v;
        } while (b > T(1) && ++turns < MAX_REJECT);
        a = T(2) * std::sqrt(T(1) - b);
        return v3<T>(a * u, a * v, T(2) * b - T(1));
    }
    // qsim::lambertian_direction (qudarap/qsim.h:430-483)
    V3<T> lambertian_direction(V3<T> normal, T orient) {
        T ndotv; int count = 0; T u; V3<T> dir;
        do {
            count++;
            dir = random_direction_marsaglia();
            ndotv = dot(dir, normal) * orient;
            if (ndotv < T(0)) { dir = T(-1) * dir; ndotv = T(-1) * ndotv; }
            u = U();
        } while (!(u < ndotv) && (count < 1024));
        return dir;
    }
    // qsim::reflect_diffuse (qudarap/qsim.h:1977-2012)
    void reflect_diffuse(Photon<T>& p, V3<T> normal) {
        V3<T> old_mom = p.mom;
        T orient = dot(old_mom, normal) > T(0) ? T(-1) : T(1);
        p.mom = lambertian_direction(normal, orient);
        V3<T> facet_normal = normalize(p.mom - old_mom);
        T EdotN = dot(p.pol, facet_normal);
        p.pol = T(-1) * p.pol + T(2) * EdotN * facet_normal;
    }
    // qsim::reflect_specular (qudarap/qsim.h:2030-2082)
    void reflect_specular(Photon<T>& p, V3<T> normal) {
        T PdotN = dot(p.mom, normal);
        p.mom = p.mom - T(2) * PdotN * normal;
        T EdotN = dot(p.pol, normal);
        p.pol = T(-1) * p.pol + T(2) * EdotN * normal;
    }
    // qsim::propagate_at_surface (qudarap/qsim.h:1677-1747), no Custom4
    int propagate_at_surface(unsigned& flag, Photon<T>& p, const State<T>& s, V3<T> normal) {
        T detect = s.surface.x, absorb = s.surface.y, reflect_diffuse_ = s.surface.w;
        T u_surface = U();
        if (aligned()) U();
        int action = u_surface < absorb + detect ? BREAK : CONTINUE;
        if (action == BREAK) {
            flag = u_surface < absorb ? SURFACE_ABSORB : SURFACE_DETECT;
        } else {
            flag = u_surface < absorb + detect + reflect_diffuse_ ? SURFACE_DREFLECT : SURFACE_SREFLECT;
            if (flag == SURFACE_DREFLECT) reflect_diffuse(p, normal);
            else reflect_specular(p, normal);
        }
        return action;
    }

    // qsim::propagate (qudarap/qsim.h:2218-2327)
    int propagate(Photon<T>& p, const PRD<T>& prd, V3<T> normal) {
        unsigned boundary = prd.gp_boundary & 0xffffu;
        unsigned identity = prd.iindex_identity & 0xffffu;
        unsigned iindex = prd.iindex_identity >> 16;
        float lposcost = prd.lposcost;
        T cosTheta = dot(p.mom, normal);
        p.set_prd(boundary, identity, cosTheta, iindex);
        State<T> s;
        fill_state(s, boundary, p.wavelength, cosTheta);
        unsigned flag = 0;
        int command = propagate_to_boundary(flag, p, s, prd.t);
        if (command == BOUNDARY) {
            int ems = int(s.optical[1]);
            if (ems == 1) command = propagate_at_boundary(flag, p, s, normal);
            else if (ems == 2) command = propagate_at_surface(flag, p, s, normal);
            else if (lposcost < 0.f) command = propagate_at_surface(flag, p, s, normal);
            else if (ems == 3) { U(); flag = SURFACE_DETECT; command = BREAK; }
        }
        p.set_flag(flag);
        return command;
    }

    // ---------------------------------------------------------- generation
    // storch::generate (sysrap/storch.h:189-516) + T_POINT_ISOTROPIC extension
    void gen_torch(Photon<T>& p, const float* gs, uint64_t photon_id) {
        uint32_t numphoton = f2u(gs[3]);
        uint32_t type = f2u(gs[23]);
        V3<T> gpos = v3<T>(gs[4], gs[5], gs[6]);
        V3<T> gmom = v3<T>(gs[8], gs[9], gs[10]);
        T wavelength = gs[15], time = gs[7];
        T zx = gs[16], zy = gs[17], ax = gs[18], ay = gs[19];
        T radius = gs[20], distance = gs[21];
        p.wavelength = wavelength; p.time = time;
        const T TWOPI = T(2) * pi<T>();
        if (type == 1) {             // T_DISC
            p.mom = gmom;
            T u_zenith = zx + U() * (zy - zx);
            T u_azimuth = ax + U() * (ay - ax);
            T r = radius * u_zenith;
            T phi = TWOPI * u_azimuth;
            T sinPhi = xsin(phi), cosPhi = xcos(phi);
            p.pos = v3<T>(r * cosPhi, r * sinPhi, T(0));
            rotateUz(p.pos, p.mom);
            p.pos = p.pos + gpos;
            p.pol = v3<T>(sinPhi, -cosPhi, T(0));
            rotateUz(p.pol, p.mom);
        } else if (type == 7) {      // T_SPHERE
            T u_zenith = zx + U() * (zy - zx);
            T u_azimuth = ax + U() * (ay - ax);
            T phi = TWOPI * u_azimuth;
            T sinPhi = xsin(phi), cosPhi = xcos(phi);
            T cosTheta = T(1) - T(2) * u_zenith;
            T sinTheta = std::sqrt(T(1) - cosTheta * cosTheta);
            T flip = std::copysign(T(1), radius);
            p.mom = v3<T>(flip * sinTheta * cosPhi, flip * sinTheta * sinPhi, flip * cosTheta);
            T rad = std::fabs(radius);
            p.pos = v3<T>(sinTheta * cosPhi * rad, sinTheta * sinPhi * rad, cosTheta * rad);
            T phase = TWOPI * distance;
            p.pol = v3<T>(xcos(phase), xsin(phase), T(0));
            rotateUz(p.pol, p.mom);
        } else if (type == 6) {      // T_SPHERE_MARSAGLIA
            T u0z, u1a, u, v, b, a;
            int turns = 0;
            do {
                u0z = zx + U() * (zy - zx);
                u1a = ax + U() * (ay - ax);
                u = T(2) * u0z - T(1);
                v = T(2) * u1a - T(1);
                b = u * u + v * v;
            } while (b > T(1) && ++turns < MAX_REJECT);
            a = T(2) * std::sqrt(T(1) - b);
            T rad = std::fabs(radius);
            T flip = std::copysign(T(1), radius);
            p.mom = v3<T>(flip * a * u, flip * a * v, flip * (T(2) * b - T(1)));
            p.pos = v3<T>(a * u * rad, a * v * rad, (T(2) * b - T(1)) * rad);
            T phase = TWOPI * distance;
            p.pol = v3<T>(xcos(phase), xsin(phase), T(0));
            rotateUz(p.pol, p.mom);
        } else if (type == 2) {      // T_LINE
            p.mom = gmom;
            T frac = T(float(photon_id)) / T(float(numphoton));
            T sfrac = T(2) * (frac - T(0.5));
            T r = radius * sfrac;
            p.pos = v3<T>(r, T(0), T(0));
            rotateUz(p.pos, p.mom);
            p.pos = p.pos + gpos;
            p.pol = v3<T>(T(0), T(-1), T(0));
            rotateUz(p.pol, p.mom);
        } else if (type == 3) {      // T_POINT
            p.mom = gmom;
            p.pos = v3<T>(0, 0, 0);
            rotateUz(p.pos, p.mom);
            p.pos = p.pos + gpos;
            p.pol = v3<T>(T(0), T(-1), T(0));
            rotateUz(p.pol, p.mom);
        } else if (type == 4) {      // T_CIRCLE
            T f = T(float(photon_id)) / T(float(numphoton));
            T frac = ax * (T(1) - f) + ay * f;
            T phi = TWOPI * frac;
            T sinPhi = xsin(phi), cosPhi = xcos(phi);
            T r = radius < T(0) ? -radius : radius;
            p.mom = v3<T>(radius < T(0) ? -cosPhi : cosPhi, T(0), radius < T(0) ? -sinPhi : sinPhi);
            p.pos = v3<T>(r * cosPhi, T(0), r * sinPhi);
            p.pos = p.pos + gpos;
            p.pol = v3<T>(T(0), T(-1), T(0));
            rotateUz(p.pol, p.mom);
        } else if (type == 5) {      // T_RECTANGLE
            int side_size = int(numphoton / 4);
            int side = int(photon_id) / side_size;
            int side_offset = side * side_size;
            int side_index = int(photon_id) - side_offset;
            T frac = T(float(side_index)) / T(float(side_size));
            if (side == 0 || side == 1) {
                p.pos = v3<T>(side == 0 ? ax : ay, T(0), (T(1) - frac) * zx + frac * zy);
                p.mom = v3<T>(side == 0 ? T(1) : T(-1), T(0), T(0));
            } else if (side == 2 || side == 3) {
                p.pos = v3<T>((T(1) - frac) * ax + frac * ay, T(0), side == 2 ? zx : zy);
                p.mom = v3<T>(T(0), T(0), side == 2 ? T(1) : T(-1));
            }
            p.pos = p.pos + gpos;
            p.pol = v3<T>(T(0), T(-1), T(0));
            rotateUz(p.pol, p.mom);
        } else if (type == 8) {      // T_POINT_ISOTROPIC (extension)
            T u0 = U(), u1 = U(), u2 = U();
            T phi = TWOPI * u0;
            T cosTheta = T(1) - T(2) * u1;
            T sinTheta = std::sqrt(T(1) - cosTheta * cosTheta);
            p.mom = v3<T>(sinTheta * xcos(phi), sinTheta * xsin(phi), cosTheta);
            p.pos = gpos;
            T phase = TWOPI * u2;
            p.pol = v3<T>(xcos(phase), xsin(phase), T(0));
            rotateUz(p.pol, p.mom);
        }
        p.zero_flags();
        p.set_flag(TORCH);
    }

    // qcerenkov::generate + wavelength_sampled_bndtex + fraction_sampled (qudarap/qcerenkov.h:56-327)
    void gen_cerenkov(Photon<T>& p, const float* gs) {
        V3<T> dp = v3<T>(gs[8], gs[9], gs[10]);
        V3<T> p0 = normalize(dp);
        unsigned matline = f2u(gs[2]);
        T BetaInverse = gs[16], Wmin = gs[17], Wmax = gs[18], maxSin2 = gs[20];
        T mean1 = gs[21], mean2 = gs[22];
        T wavelength, cosTheta, sin2Theta;
        unsigned count = 0;
        T u_maxSin2;
        do {
            T u0 = U();
            T w = Wmin + u0 * (Wmax - Wmin);
            wavelength = Wmin * Wmax / w;
            V4<T> props = boundary_lookup<T>(c, wavelength, matline, 0u);
            T sampledRI = props.x;
            cosTheta = BetaInverse / sampledRI;
            sin2Theta = fmx(T(0), (T(1) - cosTheta) * (T(1) + cosTheta));
            T u1 = U();
            u_maxSin2 = u1 * maxSin2;
            count += 1;
        } while (u_maxSin2 > sin2Theta && count < 100);
        T sinTheta = std::sqrt(sin2Theta);
        T u0 = U();
        T phi = T(2) * pi<T>() * u0;
        T sinPhi = xsin(phi), cosPhi = xcos(phi);
        p.mom = v3<T>(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
        rotateUz(p.mom, p0);
        p.pol = v3<T>(cosTheta * cosPhi, cosTheta * sinPhi, -sinTheta);
        rotateUz(p.pol, p0);
        p.wavelength = wavelength;
        T fraction, delta, N, NumberOfPhotons;
        T MeanMax = fmx(mean1, mean2);
        T DeltaN = mean1 - mean2;
        T step_length = gs[11];
        int turns = 0;
        do {
            fraction = U();
            delta = fraction * step_length;
            NumberOfPhotons = mean1 - fraction * DeltaN;
            T u = U();
            N = u * MeanMax;
        } while (N > NumberOfPhotons && ++turns < MAX_REJECT);
        T preV = gs[15], postV = gs[23];
        T midVelocity = preV + fraction * (postV - preV) * T(0.5);
        p.time = T(gs[7]) + delta / midVelocity;
        p.pos = v3<T>(gs[4], gs[5], gs[6]) + fraction * dp;
        p.zero_flags();
        p.set_flag(CERENKOV);
    }

    // qscint::generate + mom_pol_wavelength (qudarap/qscint.h:55-145)
    void gen_scint(Photon<T>& p, const float* gs) {
        T u0 = U(), u1 = U(), u2 = U(), u3 = U();
        T cost = T(1) - T(2) * u0;
        T sint = std::sqrt((T(1) - cost) * (T(1) + cost));
        T phi = T(2) * pi<T>() * u1;
        T sinp = xsin(phi), cosp = xcos(phi);
        p.mom = v3<T>(sint * cosp, sint * sinp, cost);
        p.pol = v3<T>(cost * cosp, cost * sinp, -sint);
        phi = T(2) * pi<T>() * u2;
        sinp = xsin(phi); cosp = xcos(phi);
        p.pol = normalize(cosp * p.pol + sinp * cross(p.mom, p.pol));
        p.wavelength = scint_wavelength<T>(c, u3);
        T charge = gs[13];
        T fraction = charge == T(0) ? T(1) : U();
        p.pos = v3<T>(gs[4], gs[5], gs[6]) + fraction * v3<T>(gs[8], gs[9], gs[10]);
        T u4 = U();
        T deltaTime = fraction * T(gs[11]) / T(gs[15]) - T(gs[20]) * xlog(u4);
        p.time = T(gs[7]) + deltaTime;
        p.zero_flags();
        p.set_flag(SCINTILLATION);
    }
};

struct Outputs {
    float* photon; uint64_t* seq; float* record; int32_t max_record;
    uint32_t* bounce; uint32_t* overflow;
    uint32_t* lite;     // sphotonlite (4 words) per photon or null
    uint32_t* rec; int32_t max_rec;     // srec (4 words) per (photon, bounce) slot or null
    float* prd; int32_t max_prd;        // quad2 (8 floats) per (photon, bounce) slot or null
};

// FLOAT2INT_RN (sysrap/srec.h:106-110): round to nearest even, saturating, NaN giving 0
inline int32_t f2i_rn(float x) {
    if (x != x) return 0;
    const float r = std::nearbyint(x);
    if (r >= 2147483648.f) return 2147483647;
    if (r < -2147483647.f) return INT32_MIN;
    return int32_t(r);
}

// sevent::add_rec (sysrap/sevent.h:614-622) with the srec setters (sysrap/srec.h:126-191)
// and the domains of sevent::init_domain (sevent.h:285-297): centre zero, position extent
// and time extent from the config, wavelength centre (60+820)/2 and extent (820-60)/2.
// The int results are narrowed to int16 (post) and int8 (polw) by truncation.
inline void rec_pack(uint32_t* r, float px, float py, float pz, float t, float ax, float ay, float az, float w,
                     float extent, float tdom) {
    const int32_t qx = f2i_rn((32767.0f * (px - 0.f)) / extent);
    const int32_t qy = f2i_rn((32767.0f * (py - 0.f)) / extent);
    const int32_t qz = f2i_rn((32767.0f * (pz - 0.f)) / extent);
    const int32_t qt = f2i_rn((32767.0f * (t - 0.f)) / tdom);
    const int32_t wx = f2i_rn(ax * 127.f), wy = f2i_rn(ay * 127.f), wz = f2i_rn(az * 127.f);
    const int32_t ww = f2i_rn(((w - 440.f) * 127.f) / 380.f);
    r[0] = (uint32_t(qx) & 0xffffu) | (uint32_t(qy) << 16);
    r[1] = (uint32_t(qz) & 0xffffu) | (uint32_t(qt) << 16);
    r[2] = (uint32_t(wx) & 0xffu) | ((uint32_t(wy) & 0xffu) << 8) | ((uint32_t(wz) & 0xffu) << 16) | (uint32_t(ww) << 24);
    r[3] = 0u;
}

// sphotonlite::set_lpos (sysrap/sphotonlite.h:234-243): v*0xffff+0.5 converted to 16 bits;
// the conversion saturates, negative and NaN giving 0.
inline uint32_t lite_unorm16(float v) {
    float x = v * 65535.f + 0.5f;
    if (!(x > 0.f)) return 0u;
    if (x >= 65535.f) return 65535u;
    return uint32_t(x);
}

// simulate (CSGOptiX/CSGOptiX7.cu:405-503) for one photon
template <typename T>
void simulate_one(const Ctx& c, const float* gs_all, const int32_t* seed, int64_t idx, int32_t event_index,
                  uint64_t slot_offset, const float* pre, int32_t npre, const float* input, const Outputs& out) {
    Sim<T> sim(c);
    int32_t gsi = seed[idx];
    const float* gs = gs_all + 24 * size_t(gsi);
    uint64_t photon_idx = slot_offset + uint64_t(idx);
    sim.rng.init(c.cfg.rng_seed, c.cfg.rng_offset, c.cfg.skipahead_event, uint64_t(event_index), photon_idx);
    if (pre) { sim.rng.pre = pre + size_t(idx) * npre; sim.rng.npre = npre; }
    Photon<T> p{};
    uint32_t gencode = f2u(gs[0]);
    switch (gencode) {
        case 6: sim.gen_torch(p, gs, photon_idx); break;
        case 15: case 18: sim.gen_cerenkov(p, gs); break;
        case 16: case 5: sim.gen_scint(p, gs); break;
        case 19: p.load(input + 16 * size_t(photon_idx)); p.set_flag(TORCH); break;
        default: p.zero_flags(); break;
    }
    p.set_index(photon_idx);
    uint64_t seqhis[2] = {0, 0}, seqbnd[2] = {0, 0};
    auto point = [&](int bounce) {
        if (out.record && bounce < out.max_record) p.store(out.record + 16 * (size_t(idx) * out.max_record + bounce));
        if (out.rec && bounce < out.max_rec)
            rec_pack(out.rec + 4 * (size_t(idx) * out.max_rec + bounce), float(p.pos.x), float(p.pos.y), float(p.pos.z),
                     float(p.time), float(p.pol.x), float(p.pol.y), float(p.pol.z), float(p.wavelength),
                     c.cfg.extent_domain, c.cfg.time_domain);
        if (out.seq) {
            unsigned slot = unsigned(bounce);
            unsigned iseq = slot / 16;
            unsigned shift = 4 * (slot - iseq * 16);
            if (iseq < 2) {
                unsigned fl = p.flag();
                seqhis[iseq] |= (uint64_t(unsigned(__builtin_ffs(int(fl)))) & 0xfull) << shift;
                seqbnd[iseq] |= (uint64_t(p.boundary()) & 0xfull) << shift;
            }
        }
    };
    int bounce = 0;
    float last_lposcost = 0.f, last_lposfphi = 0.f;   // PRD of the last trace (zero on a miss)
    point(bounce);
    while (bounce < c.cfg.max_bounce && p.time < T(c.cfg.max_time)) {
        T tmin = (p.orient_boundary_flag & c.cfg.eps0_mask) ? T(c.cfg.tmin0) : T(c.cfg.tmin);
        PRD<T> prd;
        trace_refined<T>(c, p.pos, p.mom, tmin, T(c.cfg.tmax), prd);
        last_lposcost = prd.lposcost; last_lposfphi = prd.lposfphi;
        if ((prd.gp_boundary & 0xffffu) == 0xffffu) break;
        V3<T> normal = normalize(prd.normal);
        if (out.prd && bounce < out.max_prd) {
            // sctx::trace (sysrap/sctx.h:170-173): quad2 of sysrap/squad.h:95-112
            float* q = out.prd + 8 * (size_t(idx) * out.max_prd + bounce);
            q[0] = float(normal.x); q[1] = float(normal.y); q[2] = float(normal.z); q[3] = float(prd.t);
            q[4] = prd.lposcost; q[5] = prd.lposfphi; q[6] = u2f(prd.iindex_identity); q[7] = u2f(prd.gp_boundary);
        }
        int command = sim.propagate(p, prd, normal);
        bounce++;
        point(bounce);
        if (command == BREAK) break;
    }
    if (out.seq) {
        uint64_t* s = out.seq + 4 * size_t(idx);
        s[0] = seqhis[0]; s[1] = seqhis[1]; s[2] = seqbnd[0]; s[3] = seqbnd[1];
    }
    p.store(out.photon + 16 * size_t(idx));
    if (out.lite) {
        // CSGOptiX7.cu:493-499: sphotonlite::init(identity, time, flagmask) + set_lpos
        uint32_t* l = out.lite + 4 * size_t(idx);
        l[0] = 0x10000u | (p.identity & 0xffffu);
        l[1] = f2u(float(p.time));
        l[2] = (lite_unorm16(last_lposcost) << 16) | lite_unorm16(last_lposfphi);
        l[3] = p.flagmask;
    }
    if (out.bounce) out.bounce[idx] = uint32_t(bounce);
    if (out.overflow) out.overflow[idx] = sim.rng.overflow;
}

} // namespace orc

// ============================================================== C entry points (tests only)
extern "C" {

void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    orc::philox4x32_10(ctr, key, out);
}

void orc_uniforms(uint64_t seed, uint64_t offset, uint64_t skip, uint64_t event, uint64_t photon, int64_t n, float* out) {
    orc::Stream s;
    s.init(seed, offset, skip, event, photon);
    for (int64_t i = 0; i < n; i++) out[i] = s.next();
}

void* orc_create(const oxr_geometry* g, const oxr_physics* p, const oxr_config* cfg) {
    auto* c = new orc::Ctx();
    c->solid = g->solid; c->num_solid = g->num_solid;
    c->prim = g->prim; c->num_prim = g->num_prim;
    c->node = reinterpret_cast<const orc::Node*>(g->node); c->num_node = g->num_node;
    c->plan = g->plan; c->num_plan = g->num_plan;
    c->itra = g->itra; c->num_itra = g->num_itra;
    c->inst = g->inst; c->num_inst = g->num_inst;
    c->w2o.resize(size_t(g->num_inst));
    c->inst_identity_transform.resize(size_t(g->num_inst));
    for (int64_t i = 0; i < g->num_inst; i++) {
        const float* q = g->inst + 16 * i;
        c->w2o[i] = orc::invert_instance(q);
        bool ident = q[0] == 1.f && q[1] == 0.f && q[2] == 0.f && q[4] == 0.f && q[5] == 1.f && q[6] == 0.f &&
                     q[8] == 0.f && q[9] == 0.f && q[10] == 1.f && q[12] == 0.f && q[13] == 0.f && q[14] == 0.f;
        c->inst_identity_transform[i] = ident ? 1 : 0;
    }
    c->bnd = p->bnd; c->num_bnd = p->num_bnd; c->num_wl = p->num_wl; c->wl_min = p->wl_min; c->wl_step = p->wl_step;
    c->optical = p->optical;
    c->icdf = p->icdf; c->icdf_rows = p->icdf_rows; c->icdf_width = p->icdf_width; c->hd_factor = p->hd_factor;
    c->cfg = *cfg;
    return c;
}

void orc_destroy(void* c) { delete static_cast<orc::Ctx*>(c); }

void orc_set_config(void* c, const oxr_config* cfg) { static_cast<orc::Ctx*>(c)->cfg = *cfg; }

// enable conservative box culling in trace (instance world boxes, prim object boxes)
void orc_set_cull(void* cv, int32_t on) {
    orc::Ctx& c = *static_cast<orc::Ctx*>(cv);
    c.cull = on;
    if (!on || !c.inst_box.empty()) return;
    c.prim_box.resize(6 * size_t(c.num_prim));
    for (int64_t p = 0; p < c.num_prim; p++) {
        const float* q = c.prim + 16 * p;
        for (int k = 0; k < 6; k++) c.prim_box[6 * p + k] = q[8 + k];
    }
    c.inst_box.resize(6 * size_t(c.num_inst));
    for (int64_t i = 0; i < c.num_inst; i++) {
        const float* q = c.inst + 16 * i;
        int32_t gas; std::memcpy(&gas, q + 7, 4);
        const int32_t* so = c.solid + 12 * size_t(gas);
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int p = 0; p < so[4]; p++) {
            const double* b = &c.prim_box[6 * size_t(so[5] + p)];
            for (int cc = 0; cc < 8; cc++) {
                double pt[3] = {(cc & 1) ? b[3] : b[0], (cc & 2) ? b[4] : b[1], (cc & 4) ? b[5] : b[2]};
                for (int k = 0; k < 3; k++) {
                    double w = pt[0] * q[k] + pt[1] * q[4 + k] + pt[2] * q[8 + k] + q[12 + k];
                    lo[k] = std::min(lo[k], w); hi[k] = std::max(hi[k], w);
                }
            }
        }
        for (int k = 0; k < 3; k++) { c.inst_box[6 * i + k] = lo[k]; c.inst_box[6 * i + 3 + k] = hi[k]; }
    }
}

// seed expansion: QEvt_count_genstep_photons_and_fill_seed_buffer (qudarap/QEvt.cu:181-237)
int64_t orc_fill_seed(const float* gs, int64_t num_gs, int32_t* seed, int64_t cap) {
    int64_t n = 0;
    for (int64_t g = 0; g < num_gs; g++) {
        uint32_t np = orc::f2u(gs[24 * g + 3]);
        for (uint32_t k = 0; k < np; k++) { if (seed && n < cap) seed[n] = int32_t(g); n++; }
    }
    return n;
}

// intersect_prim in the prim's frame; precision 0 float, 1 double
void orc_intersect_prim(void* cv, int32_t prim_index, const float* rays, int64_t n, float* isect, int32_t* valid,
                        int32_t precision) {
    const orc::Ctx& c = *static_cast<orc::Ctx*>(cv);
    const float* pr = c.prim + 16 * size_t(prim_index);
    int32_t nodeOffset; std::memcpy(&nodeOffset, pr + 1, 4);
    const orc::Node* root = c.node + nodeOffset;
    for (int64_t i = 0; i < n; i++) {
        const float* r = rays + 8 * i;
        if (precision == 0) {
            orc::Isect<float> is;
            bool v = orc::intersect_prim<float>(c, is, root, r[3], orc::v3<float>(r[0], r[1], r[2]), orc::v3<float>(r[4], r[5], r[6]));
            isect[4 * i + 0] = is.x; isect[4 * i + 1] = is.y; isect[4 * i + 2] = is.z; isect[4 * i + 3] = is.w;
            valid[i] = v;
        } else {
            orc::Isect<double> is;
            bool v = orc::intersect_prim<double>(c, is, root, r[3], orc::v3<double>(r[0], r[1], r[2]), orc::v3<double>(r[4], r[5], r[6]));
            isect[4 * i + 0] = float(is.x); isect[4 * i + 1] = float(is.y); isect[4 * i + 2] = float(is.z); isect[4 * i + 3] = float(is.w);
            valid[i] = v;
        }
    }
}

// closest hit over all instances x prims: prd (n,2,4)
void orc_trace_rays(void* cv, const float* rays, int64_t n, float* prd, int32_t precision) {
    const orc::Ctx& c = *static_cast<orc::Ctx*>(cv);
    for (int64_t i = 0; i < n; i++) {
        const float* r = rays + 8 * i;
        float* o = prd + 8 * i;
        if (precision == 0) {
            orc::PRD<float> q;
            orc::trace<float>(c, orc::v3<float>(r[0], r[1], r[2]), orc::v3<float>(r[4], r[5], r[6]), r[3], r[7], q);
            o[0] = q.normal.x; o[1] = q.normal.y; o[2] = q.normal.z; o[3] = q.t;
            o[4] = q.lposcost; o[5] = q.lposfphi; o[6] = orc::u2f(q.iindex_identity); o[7] = orc::u2f(q.gp_boundary);
        } else {
            orc::PRD<double> q;
            orc::trace<double>(c, orc::v3<double>(r[0], r[1], r[2]), orc::v3<double>(r[4], r[5], r[6]), r[3], r[7], q);
            o[0] = float(q.normal.x); o[1] = float(q.normal.y); o[2] = float(q.normal.z); o[3] = float(q.t);
            o[4] = q.lposcost; o[5] = q.lposfphi; o[6] = orc::u2f(q.iindex_identity); o[7] = orc::u2f(q.gp_boundary);
        }
    }
}

// single-precision log of the photon loop, for its accuracy pin
void orc_xlog(const float* x, int64_t n, float* out) {
    for (int64_t i = 0; i < n; i++) out[i] = orc::xlog<float>(x[i]);
}

// simtrace over FRAME gensteps: qsim::generate_photon_simtrace_frame (qudarap/qsim.h:2459-2510)
// then trace<refine> (CSGOptiX7.cu:536-577) and sevent::add_simtrace (sysrap/sevent.h:670-696);
// out (n,4,4). Returns n.
int64_t orc_simtrace(void* cv, const float* gs, int64_t num_gs, float* out, int64_t cap) {
    const orc::Ctx& c = *static_cast<orc::Ctx*>(cv);
    int64_t total = orc_fill_seed(gs, num_gs, nullptr, 0);
    if (!out || total > cap) return total;
    std::vector<int32_t> seed(static_cast<size_t>(total));
    orc_fill_seed(gs, num_gs, seed.data(), total);
    for (int64_t i = 0; i < total; i++) {
        const float* q = gs + 24 * size_t(seed[size_t(i)]);
        const uint32_t gridaxes = orc::f2u(q[1]);
        orc::Stream rng;
        rng.init(c.cfg.rng_seed, 0, 0, 0, uint64_t(i));
        float u0 = rng.next();
        float ph = 2.f * orc::pi<float>() * u0;
        float sinPhi = orc::xsin(ph), cosPhi = orc::xcos(ph);
        float u1 = rng.next();
        float cosTheta = 2.f * u1 - 1.f;
        float sinTheta = std::sqrt(1.f - cosTheta * cosTheta);
        float d[3];
        switch (gridaxes) {
            case 1: d[0] = 0.f; d[1] = cosPhi; d[2] = sinPhi; break;
            case 2: d[0] = cosPhi; d[1] = 0.f; d[2] = sinPhi; break;
            case 3: d[0] = cosPhi; d[1] = sinPhi; d[2] = 0.f; break;
            default: d[0] = sinTheta * cosPhi; d[1] = sinTheta * sinPhi; d[2] = cosTheta; break;
        }
        const float* m = q + 8;          // qat4 rows q2..q5, translation in the last
        const float p[3] = {q[4], q[5], q[6]};
        float o[3], v[3];
        for (int k = 0; k < 3; k++) {
            o[k] = m[k] * p[0] + m[4 + k] * p[1] + m[8 + k] * p[2] + m[12 + k] * 1.f;
            v[k] = m[k] * d[0] + m[4 + k] * d[1] + m[8 + k] * d[2] + m[12 + k] * 0.f;
        }
        orc::PRD<float> r;
        const bool found = orc::trace_refined<float>(c, orc::v3<float>(o[0], o[1], o[2]), orc::v3<float>(v[0], v[1], v[2]),
                                                     c.cfg.tmin, c.cfg.tmax, r);
        // a miss carries the miss program's q0: mid-grey background normal, distance 1
        // (CSGOptiX7.cu:655-682, SBT::updateMiss SBT.cc:180-192)
        if (!found) r.normal = orc::v3<float>(0.6f, 0.6f, 0.6f);
        const float t = r.t;
        float* a = out + 16 * size_t(i);
        a[0] = r.normal.x; a[1] = r.normal.y; a[2] = r.normal.z; a[3] = t;
        a[4] = o[0] + t * v[0]; a[5] = o[1] + t * v[1]; a[6] = o[2] + t * v[2]; a[7] = c.cfg.tmin;
        a[8] = o[0]; a[9] = o[1]; a[10] = o[2]; a[11] = orc::u2f(r.gp_boundary);
        a[12] = v[0]; a[13] = v[1]; a[14] = v[2]; a[15] = orc::u2f(r.iindex_identity);
    }
    return total;
}

// Full simulate of one event slice. seq (n,4) u64 may be null; record may be null.
// photon_begin/photon_end select a sub-range of the slice (for sampling); nthreads std::threads.
int64_t orc_simulate(void* cv, const float* gs, int64_t num_gs, int32_t event_index, uint64_t slot_offset,
                     const float* precooked, int32_t npre, const float* input_photons,
                     int32_t precision, int32_t nthreads, int64_t photon_begin, int64_t photon_end,
                     float* photon, uint64_t* seq, float* record, int32_t max_record, uint32_t* bounce,
                     uint32_t* overflow, uint32_t* lite, uint32_t* rec, int32_t max_rec, float* prd,
                     int32_t max_prd) {
    const orc::Ctx& c = *static_cast<orc::Ctx*>(cv);
    int64_t total = orc_fill_seed(gs, num_gs, nullptr, 0);
    std::vector<int32_t> seed(static_cast<size_t>(total));
    orc_fill_seed(gs, num_gs, seed.data(), total);
    if (photon_end < 0 || photon_end > total) photon_end = total;
    if (photon_begin < 0) photon_begin = 0;
    orc::Outputs out{photon, seq, record, max_record, bounce, overflow, lite, rec, max_rec, prd, max_prd};
    std::atomic<int64_t> next(photon_begin);
    auto work = [&]() {
        for (;;) {
            int64_t i0 = next.fetch_add(256);
            if (i0 >= photon_end) break;
            int64_t i1 = std::min(photon_end, i0 + 256);
            for (int64_t i = i0; i < i1; i++) {
                if (precision == 0)
                    orc::simulate_one<float>(c, gs, seed.data(), i, event_index, slot_offset, precooked, npre, input_photons, out);
                else
                    orc::simulate_one<double>(c, gs, seed.data(), i, event_index, slot_offset, precooked, npre, input_photons, out);
            }
        }
    };
    if (nthreads <= 1) work();
    else {
        std::vector<std::thread> th;
        for (int t = 0; t < nthreads; t++) th.emplace_back(work);
        for (auto& t : th) t.join();
    }
    return total;
}

// per-function physics entry points for closed-form checks (tests only)
// propagate_at_boundary on given mom/pol/normal with n1,n2 and a precooked stream
void orc_propagate_at_boundary(const float* mom_pol_normal, float n1, float n2, const float* u, int32_t nu,
                               int32_t n, float* out_mom_pol_flag) {
    orc::Ctx c{};
    c.cfg.rng_schedule = OXR_RNG_PRODUCTION;
    for (int32_t i = 0; i < n; i++) {
        orc::Sim<float> sim(c);
        sim.rng.init(0, 0, 0, 0, 0);
        sim.rng.pre = u + size_t(i) * nu; sim.rng.npre = nu;
        orc::Photon<float> p{};
        const float* q = mom_pol_normal + 9 * size_t(i);
        p.mom = orc::v3<float>(q[0], q[1], q[2]);
        p.pol = orc::v3<float>(q[3], q[4], q[5]);
        orc::State<float> s{};
        s.material1.x = n1; s.material2.x = n2;
        unsigned flag = 0;
        sim.propagate_at_boundary(flag, p, s, orc::v3<float>(q[6], q[7], q[8]));
        float* o = out_mom_pol_flag + 7 * size_t(i);
        o[0] = p.mom.x; o[1] = p.mom.y; o[2] = p.mom.z; o[3] = p.pol.x; o[4] = p.pol.y; o[5] = p.pol.z;
        o[6] = float(flag);
    }
}

// Rayleigh scatter / Lambertian / Marsaglia sampling distributions from a Philox stream
void orc_sample(int32_t which, const float* mom_pol_normal, int64_t n, uint64_t seed, float* out) {
    orc::Ctx c{};
    for (int64_t i = 0; i < n; i++) {
        orc::Sim<float> sim(c);
        sim.rng.init(seed, 0, 0, 0, uint64_t(i));
        orc::Photon<float> p{};
        p.mom = orc::v3<float>(mom_pol_normal[0], mom_pol_normal[1], mom_pol_normal[2]);
        p.pol = orc::v3<float>(mom_pol_normal[3], mom_pol_normal[4], mom_pol_normal[5]);
        orc::V3<float> nrm = orc::v3<float>(mom_pol_normal[6], mom_pol_normal[7], mom_pol_normal[8]);
        if (which == 0) sim.rayleigh_scatter(p);
        else if (which == 1) sim.reflect_diffuse(p, nrm);
        else if (which == 2) p.mom = sim.random_direction_marsaglia();
        else if (which == 3) sim.reflect_specular(p, nrm);
        float* o = out + 6 * i;
        o[0] = p.mom.x; o[1] = p.mom.y; o[2] = p.mom.z; o[3] = p.pol.x; o[4] = p.pol.y; o[5] = p.pol.z;
    }
}

// boundary_lookup for table checks
void orc_boundary_lookup(void* cv, const float* nm, int64_t n, uint32_t line, uint32_t k, float* out) {
    const orc::Ctx& c = *static_cast<orc::Ctx*>(cv);
    for (int64_t i = 0; i < n; i++) {
        orc::V4<float> v = orc::boundary_lookup<float>(c, nm[i], line, k);
        out[4 * i] = v.x; out[4 * i + 1] = v.y; out[4 * i + 2] = v.z; out[4 * i + 3] = v.w;
    }
}

void orc_scint_wavelength(void* cv, const float* u, int64_t n, float* out) {
    const orc::Ctx& c = *static_cast<orc::Ctx*>(cv);
    for (int64_t i = 0; i < n; i++) out[i] = orc::scint_wavelength<float>(c, u[i]);
}

int32_t orc_lut(int32_t op, int32_t a, int32_t b, int32_t acloser) { return orc::lut(op, a, b, acloser != 0); }

} // extern "C"
