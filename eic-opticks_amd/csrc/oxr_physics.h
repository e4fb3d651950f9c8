// oxr_physics.h -- photon generation and propagation on the device: the qsim
// generate + propagate path (qudarap/qsim.h, qbnd.h, qscint.h, qcerenkov.h and
// sysrap/storch.h) restated for gfx950 registers.
#pragma once
#include "oxr_trace.h"

// Generation runs once per photon; compiling it out of line keeps the bounce loop's code
// smaller. OXR_GEN_INLINE=1 inlines it instead.
#ifndef OXR_GEN_INLINE
#define OXR_GEN_INLINE 1
#endif
#if OXR_GEN_INLINE
#define OXG __device__ __forceinline__
#else
#define OXG __device__ __noinline__
#endif

namespace oxd {

struct Ph {
    f3 pos; float time;
    f3 mom; uint32_t hitcount_iindex;
    f3 pol; float wavelength;
    uint32_t obf, identity, index, flagmask;
};

OXD unsigned ph_flag(const Ph& p) { return p.obf & 0xffffu; }
OXD unsigned ph_boundary(const Ph& p) { return (p.obf & 0x7fff0000u) >> 16; }
OXD void set_flag(Ph& p, unsigned f) { p.obf = (p.obf & 0xffff0000u) | (f & 0xffffu); p.flagmask |= f; }
OXD void zero_flags(Ph& p) { p.obf = 0; p.identity = 0; p.index = 0; p.flagmask = 0; p.hitcount_iindex = 0; }
// sphoton::set_prd (sysrap/sphoton.h:482-488)
OXD void set_prd(Ph& p, unsigned boundary, unsigned id, float orient, unsigned iindex) {
    p.obf = (p.obf & 0x8000ffffu) | ((boundary & 0x7fffu) << 16);
    p.identity = (p.identity & 0xff000000u) | (id & 0x00ffffffu);
    p.obf = (p.obf & 0x7fffffffu) | ((orient < 0.f ? 1u : 0u) << 31);
    p.hitcount_iindex = 0x00010000u | (iindex & 0xffffu);
}
OXD void set_index(Ph& p, uint64_t full) {
    p.index = uint32_t(full & 0xffffffffu);
    p.identity = (uint32_t((full >> 32) & 0xffu) << 24) | (p.identity & 0xffffffu);
}
OXD void store_photon(float4* dst, const Ph& p) {
    dst[0] = make_float4(p.pos.x, p.pos.y, p.pos.z, p.time);
    dst[1] = make_float4(p.mom.x, p.mom.y, p.mom.z, __uint_as_float(p.hitcount_iindex));
    dst[2] = make_float4(p.pol.x, p.pol.y, p.pol.z, p.wavelength);
    dst[3] = make_float4(__uint_as_float(p.obf), __uint_as_float(p.identity), __uint_as_float(p.index), __uint_as_float(p.flagmask));
}

// sphotonlite (sysrap/sphotonlite.h:47-52,184-190,234-243): hitcount 1 | identity low 16 bits,
// time, lposcost and lposfphi as 16-bit unorms, flagmask. set_lpos converts
// v*0xffff+0.5 to 16 bits; that conversion saturates (negative and NaN give 0).
OXD uint32_t unorm16(float v) {
    float x = v * 65535.f + 0.5f;
    if (!(x > 0.f)) return 0u;
    if (x >= 65535.f) return 65535u;
    return uint32_t(x);
}
OXD uint4 lite_pack(const Ph& p, float lposcost, float lposfphi) {
    return make_uint4(0x10000u | (p.identity & 0xffffu), __float_as_uint(p.time),
                      (unorm16(lposcost) << 16) | unorm16(lposfphi), p.flagmask);
}
// srec (sysrap/srec.h:39-43) as filled by sevent::add_rec (sysrap/sevent.h:614-622) with
// the domains of sevent::init_domain (sevent.h:285-297): position and time over
// centre-extent domains centred on zero as int16, polarization and wavelength as int8,
// flag word zero. FLOAT2INT_RN (srec.h:106-110) is the round-to-nearest-even conversion,
// saturating with NaN giving 0; its int result is narrowed to 16 or 8 bits by truncation,
// so values outside a domain wrap round, as the reference notes (srec.h:117-120).
OXD int f2i_rn(float x) {
    if (!(x == x)) return 0;
    const float r = rintf(x);
    if (r >= 2147483648.f) return 2147483647;
    if (r < -2147483647.f) return int(0x80000000u);
    return int(r);
}
constexpr float REC_W_CENTER = 440.f;   // sevent w_center = (60 + 820) / 2 (sysrap/sevent.h:111-114)
constexpr float REC_W_EXTENT = 380.f;   // sevent w_extent = (820 - 60) / 2
OXD uint4 rec_pack(const Ph& p, float extent, float tdom) {
    const int px = f2i_rn((32767.0f * (p.pos.x - 0.f)) / extent);
    const int py = f2i_rn((32767.0f * (p.pos.y - 0.f)) / extent);
    const int pz = f2i_rn((32767.0f * (p.pos.z - 0.f)) / extent);
    const int pt = f2i_rn((32767.0f * (p.time - 0.f)) / tdom);
    const int ax = f2i_rn(p.pol.x * 127.f), ay = f2i_rn(p.pol.y * 127.f), az = f2i_rn(p.pol.z * 127.f);
    const int aw = f2i_rn(((p.wavelength - REC_W_CENTER) * 127.f) / REC_W_EXTENT);
    return make_uint4((uint32_t(px) & 0xffffu) | (uint32_t(py) << 16), (uint32_t(pz) & 0xffffu) | (uint32_t(pt) << 16),
                      (uint32_t(ax) & 0xffu) | ((uint32_t(ay) & 0xffu) << 8) | ((uint32_t(az) & 0xffu) << 16) |
                          (uint32_t(aw) << 24),
                      0u);
}
OXD void load_photon(Ph& p, const float4* s) {
    float4 a = s[0], b = s[1], c = s[2], e = s[3];
    p.pos = mk(a.x, a.y, a.z); p.time = a.w;
    p.mom = mk(b.x, b.y, b.z); p.hitcount_iindex = __float_as_uint(b.w);
    p.pol = mk(c.x, c.y, c.z); p.wavelength = c.w;
    p.obf = __float_as_uint(e.x); p.identity = __float_as_uint(e.y); p.index = __float_as_uint(e.z); p.flagmask = __float_as_uint(e.w);
}

struct St { float4 m1, m1g2, m2, su; uint4 opt; };

// smath::rotateUz (sysrap/smath.h:77-96)
OXD void rotateUz(f3& d, f3 u) {
    float up = u.x * u.x + u.y * u.y;
    if (up > 0.f) {
        up = sqrtf(up);
        float px = d.x, py = d.y, pz = d.z;
        d.x = (u.x * u.z * px - u.y * py) / up + u.x * pz;
        d.y = (u.y * u.z * px + u.x * py) / up + u.y * pz;
        d.z = -up * px + u.z * pz;
    } else if (u.z < 0.f) {
        d.x = -d.x; d.z = -d.z;
    }
}

// Rejection loops that the reference leaves unbounded stop after this many turns, so a
// degenerate precooked stream cannot spin a lane forever. With Philox uniforms each turn
// accepts with probability >= 1/4, so the cap is never reached in practice.
constexpr int MAX_REJECT = 10000;

// qbnd::fill_state (qudarap/qbnd.h:184-214)
OXD void fill_state(const Tables& t, St& s, unsigned boundary, float wl, float cosTheta) {
    int line = int(boundary) * 4;
    int m1 = cosTheta > 0.f ? line + 3 : line + 0;
    int m2 = cosTheta > 0.f ? line + 0 : line + 3;
    int su = cosTheta > 0.f ? line + 2 : line + 1;
    const TexX X = tex_x(t, wl);   // one wavelength coordinate for the four rows
    s.m1 = boundary_fetch(t, X, m1, 0);
    s.m1g2 = boundary_fetch(t, X, m1, 1);
    s.m2 = boundary_fetch(t, X, m2, 0);
    s.su = boundary_fetch(t, X, su, 0);
    s.opt = t.optical[su];
}

// qsim::uniform_sphere (qudarap/qsim.h:177-183)
OXD f3 uniform_sphere(float u0, float u1) {
    float phi = u0 * 2.f * PIf;
    float cosTheta = 2.f * u1 - 1.f;
    float sinTheta = sqrtf(1.f - cosTheta * cosTheta);
    return mk(xcos(phi) * sinTheta, xsin(phi) * sinTheta, cosTheta);
}

// qsim::rayleigh_scatter (qudarap/qsim.h:601-689)
OXD void rayleigh_scatter(Ph& p, Rng& r) {
    f3 direction, polarization;
    bool looping = true;
    int turns = 0;
    do {
        float u0 = uniform(r), u1 = uniform(r), u2 = uniform(r), u3 = uniform(r), u4 = uniform(r);
        float cosTheta = u0;
        float sinTheta = sqrtf(1.0f - u0 * u0);
        if (u1 < 0.5f) cosTheta = -cosTheta;
        float ph = 2.f * PIf * u2;
        float sinPhi = xsin(ph), cosPhi = xcos(ph);
        direction = mk(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
        rotateUz(direction, p.mom);
        float constant = -dot(direction, p.pol);
        polarization = mk(p.pol.x + constant * direction.x, p.pol.y + constant * direction.y,
                          p.pol.z + constant * direction.z);
        if (dot(polarization, polarization) == 0.f) {
            float ph3 = 2.f * PIf * u3;
            sinPhi = xsin(ph3); cosPhi = xcos(ph3);
            polarization = mk(cosPhi, sinPhi, 0.f);
            rotateUz(polarization, direction);
        } else {
            if (u3 < 0.5f) polarization = neg(polarization);
        }
        polarization = normalize(polarization);
        float doCosTheta = dot(polarization, p.pol);
        float doCosTheta2 = doCosTheta * doCosTheta;
        looping = doCosTheta2 < u4;
    } while (looping && ++turns < MAX_REJECT);
    p.mom = direction;
    p.pol = polarization;
}

// qsim::propagate_to_boundary (qudarap/qsim.h:718-863)
OXD int propagate_to_boundary(unsigned& flag, Ph& p, const St& s, float dist, Rng& r, const Tables& t, int aligned) {
    float absorption_length = s.m1.y, scattering_length = s.m1.z, reemission_prob = s.m1.w;
    float group_velocity = s.m1g2.x;
    if (aligned) { uniform(r); uniform(r); }
    float u_scattering = uniform(r);
    float u_absorption = uniform(r);
    float scattering_distance = -scattering_length * xlog(u_scattering);
    float absorption_distance = -absorption_length * xlog(u_absorption);
    if (absorption_distance <= scattering_distance) {
        if (absorption_distance <= dist) {
            p.time += absorption_distance / group_velocity;
            p.pos = add(p.pos, mul(absorption_distance, p.mom));
            float u_reemit = reemission_prob == 0.f ? 2.f : uniform(r);
            if (u_reemit < reemission_prob) {
                float u_wl = uniform(r), u_mph = uniform(r), u_mct = uniform(r), u_pph = uniform(r), u_pct = uniform(r);
                p.wavelength = scint_wavelength(t, u_wl);
                p.mom = uniform_sphere(u_mph, u_mct);
                p.pol = normalize(cross(uniform_sphere(u_pph, u_pct), p.mom));
                flag = BULK_REEMIT;
                return CONTINUE;
            }
            flag = BULK_ABSORB;
            return BREAK;
        }
    } else {
        if (scattering_distance <= dist) {
            p.time += scattering_distance / group_velocity;
            p.pos = add(p.pos, mul(scattering_distance, p.mom));
            rayleigh_scatter(p, r);
            flag = BULK_SCATTER;
            return CONTINUE;
        }
    }
    p.pos = add(p.pos, mul(dist, p.mom));
    p.time += dist / group_velocity;
    return BOUNDARY;
}

// qsim::propagate_at_boundary (qudarap/qsim.h:998-1205)
OXD int propagate_at_boundary(unsigned& flag, Ph& p, const St& s, f3 normal, Rng& r, int aligned) {
    float n1 = s.m1.x, n2 = s.m2.x;
    float eta = n1 / n2;
    float _c1 = -dot(p.mom, normal);
    f3 on = _c1 < 0.f ? neg(normal) : normal;
    f3 trans = cross(p.mom, on);
    float trans_length = len(trans);
    bool normal_incidence = trans_length < 1e-6f;
    f3 A_trans = normal_incidence ? p.pol : divs(trans, trans_length);
    float E1_perp = dot(p.pol, A_trans);
    float c1 = fabsf(_c1);
    float c2c2 = 1.f - eta * eta * (1.f - c1 * c1);
    bool tir = c2c2 < 0.f;
    float EdotN = dot(p.pol, on);
    float c2 = tir ? 0.f : sqrtf(c2c2);
    float n1c1 = n1 * c1, n2c2 = n2 * c2, n2c1 = n2 * c1, n1c2 = n1 * c2;
    float E1x = normal_incidence ? 0.f : E1_perp;
    float E1y = normal_incidence ? 1.f : len(sub(p.pol, mul(E1_perp, A_trans)));
    float E2tx = 2.f * n1c1 * E1x / (n1c1 + n2c2);
    float E2ty = 2.f * n1c1 * E1y / (n2c1 + n1c2);
    float E2rx = E2tx - E1x;
    float E2ry = (n2 * E2ty / n1) - E1y;
    float rinv = 1.0f / sqrtf(E2rx * E2rx + E2ry * E2ry);
    float RRx = E2rx * rinv, RRy = E2ry * rinv;
    float tinv = 1.0f / sqrtf(E2tx * E2tx + E2ty * E2ty);
    float TTx = E2tx * tinv, TTy = E2ty * tinv;
    float TransCoeff = (tir || n1c1 == 0.f) ? 0.f : n2c2 * (E2tx * E2tx + E2ty * E2ty) / n1c1;
    if (aligned) uniform(r);
    float u_reflect = uniform(r);
    bool reflect = u_reflect > TransCoeff;
    p.mom = reflect ? add(p.mom, mul(2.f * c1, on)) : add(mul(eta, p.mom), mul(eta * c1 - c2, on));
    f3 A_paral = normalize(cross(p.mom, A_trans));
    if (normal_incidence) {
        p.pol = reflect ? mulr(p.pol, n2 > n1 ? -1.f : 1.f) : p.pol;
    } else if (reflect) {
        p.pol = tir ? add(neg(p.pol), mul(2.f * EdotN, on)) : add(mul(RRx, A_trans), mul(RRy, A_paral));
    } else {
        p.pol = add(mul(TTx, A_trans), mul(TTy, A_paral));
    }
    flag = reflect ? BOUNDARY_REFLECT : BOUNDARY_TRANSMIT;
    if (aligned && reflect) { uniform(r); uniform(r); uniform(r); uniform(r); }
    return CONTINUE;
}

// qsim::random_direction_marsaglia (qudarap/qsim.h:551-572)
OXD f3 random_direction_marsaglia(Rng& r) {
    float u, v, b, a;
    int turns = 0;
    do {
        float u0 = uniform(r), u1 = uniform(r);
        u = 2.f * u0 - 1.f;
        v = 2.f * u1 - 1.f;
        b = u * u + v * v;
    } while (b > 1.f && ++turns < MAX_REJECT);
    a = 2.f * sqrtf(1.f - b);
    return mk(a * u, a * v, 2.f * b - 1.f);
}
// qsim::lambertian_direction (qudarap/qsim.h:430-483) with random_direction_marsaglia
// (qsim.h:551-572) inlined. The reference nests the Marsaglia rejection loop inside the
// Lambertian acceptance loop; here both run as one loop, one Marsaglia try per turn, so
// the lanes of a wave advance through their own tries independently instead of every
// outer turn waiting for the slowest lane's inner loop. Each lane draws the same uniforms
// in the same order, and both turn caps (MAX_REJECT Marsaglia tries per direction, 1024
// directions) end the loop where the nested form ends.
OXD f3 lambertian_direction(f3 normal, float orient, Rng& r) {
    int count = 0, turns = 0;
    f3 dir = mk(0.f, 0.f, 0.f);
    for (;;) {
        float u0 = uniform(r), u1 = uniform(r);
        float mu = 2.f * u0 - 1.f;
        float mv = 2.f * u1 - 1.f;
        float b = mu * mu + mv * mv;
        if (b > 1.f && ++turns < MAX_REJECT) continue;        // Marsaglia: try again
        turns = 0;
        count++;
        float a = 2.f * sqrtf(1.f - b);
        dir = mk(a * mu, a * mv, 2.f * b - 1.f);
        float ndotv = dot(dir, normal) * orient;
        if (ndotv < 0.f) { dir = mul(-1.f, dir); ndotv = -1.f * ndotv; }
        float u = uniform(r);
        if (u < ndotv || count >= 1024) break;
    }
    return dir;
}
// qsim::reflect_diffuse (qudarap/qsim.h:1977-2012)
OXD void reflect_diffuse(Ph& p, f3 normal, Rng& r) {
    rng_topup(r);
    f3 old_mom = p.mom;
    float orient = dot(old_mom, normal) > 0.f ? -1.f : 1.f;
    p.mom = lambertian_direction(normal, orient, r);
    f3 fn = normalize(sub(p.mom, old_mom));
    float EdotN = dot(p.pol, fn);
    p.pol = add(mul(-1.f, p.pol), mul(2.f * EdotN, fn));
}
// qsim::reflect_specular (qudarap/qsim.h:2030-2082)
OXD void reflect_specular(Ph& p, f3 normal) {
    float PdotN = dot(p.mom, normal);
    p.mom = sub(p.mom, mul(2.f * PdotN, normal));
    float EdotN = dot(p.pol, normal);
    p.pol = add(mul(-1.f, p.pol), mul(2.f * EdotN, normal));
}
// qsim::propagate_at_surface (qudarap/qsim.h:1677-1747)
OXD int propagate_at_surface(unsigned& flag, Ph& p, const St& s, f3 normal, Rng& r, int aligned) {
    float detect = s.su.x, absorb = s.su.y, rdiff = s.su.w;
    float u_surface = uniform(r);
    if (aligned) uniform(r);
    int action = u_surface < absorb + detect ? BREAK : CONTINUE;
    if (action == BREAK) {
        flag = u_surface < absorb ? SURFACE_ABSORB : SURFACE_DETECT;
    } else {
        flag = u_surface < absorb + detect + rdiff ? SURFACE_DREFLECT : SURFACE_SREFLECT;
        if (flag == SURFACE_DREFLECT) reflect_diffuse(p, normal, r);
        else reflect_specular(p, normal);
    }
    return action;
}

// qsim::propagate (qudarap/qsim.h:2218-2327)
// normalize_cost (sysrap/scuda.h): cos(theta) of the object-frame hit position
OXD float lpos_cost(f3 lpos) { return lpos.z / sqrtf(dot(lpos, lpos)); }

// The hit's lposcost is formed only on the branch that reads it.
OXD int propagate(Ph& p, const Hit& h, f3 normal, Rng& r, const Tables& t, int aligned) {
    unsigned boundary = h.gpbd & 0xffffu;
    unsigned identity = h.ii_id & 0xffffu;
    unsigned iindex = h.ii_id >> 16;
    float cosTheta = dot(p.mom, normal);
    set_prd(p, boundary, identity, cosTheta, iindex);
    St s;
    fill_state(t, s, boundary, p.wavelength, cosTheta);
    rng_topup(r);
    unsigned flag = 0;
    int command = propagate_to_boundary(flag, p, s, h.t, r, t, aligned);
    if (command == BOUNDARY) {
        int ems = int(s.opt.y);
        if (ems == 1) command = propagate_at_boundary(flag, p, s, normal, r, aligned);
        else if (ems == 2) command = propagate_at_surface(flag, p, s, normal, r, aligned);
        else if (lpos_cost(h.lpos) < 0.f) command = propagate_at_surface(flag, p, s, normal, r, aligned);
        else if (ems == 3) { uniform(r); flag = SURFACE_DETECT; command = BREAK; }
    }
    set_flag(p, flag);
    return command;
}

// ------------------------------------------------------------------ generation
// storch::generate (sysrap/storch.h:189-516) plus the T_POINT_ISOTROPIC (8) extension
OXG void gen_torch(Ph& p, const float4* gs, uint64_t photon_id, Rng& r) {
    float4 q0 = gs[0], q1 = gs[1], q2 = gs[2], q3 = gs[3], q4 = gs[4], q5 = gs[5];
    uint32_t numphoton = __float_as_uint(q0.w);
    uint32_t type = __float_as_uint(q5.w);
    f3 gpos = mk(q1.x, q1.y, q1.z), gmom = mk(q2.x, q2.y, q2.z);
    float zx = q4.x, zy = q4.y, ax = q4.z, ay = q4.w, radius = q5.x, distance = q5.y;
    p.wavelength = q3.w; p.time = q1.w;
    const float TWOPI = 2.f * PIf;
    if (type == 1) {
        p.mom = gmom;
        float u_zenith = zx + uniform(r) * (zy - zx);
        float u_azimuth = ax + uniform(r) * (ay - ax);
        float rr = radius * u_zenith;
        float phi = TWOPI * u_azimuth;
        float sinPhi = xsin(phi), cosPhi = xcos(phi);
        p.pos = mk(rr * cosPhi, rr * sinPhi, 0.f);
        rotateUz(p.pos, p.mom);
        p.pos = add(p.pos, gpos);
        p.pol = mk(sinPhi, -cosPhi, 0.f);
        rotateUz(p.pol, p.mom);
    } else if (type == 7) {
        float u_zenith = zx + uniform(r) * (zy - zx);
        float u_azimuth = ax + uniform(r) * (ay - ax);
        float phi = TWOPI * u_azimuth;
        float sinPhi = xsin(phi), cosPhi = xcos(phi);
        float cosTheta = 1.f - 2.f * u_zenith;
        float sinTheta = sqrtf(1.f - cosTheta * cosTheta);
        float flip = copysignf(1.f, radius);
        p.mom = mk(flip * sinTheta * cosPhi, flip * sinTheta * sinPhi, flip * cosTheta);
        float rad = fabsf(radius);
        p.pos = mk(sinTheta * cosPhi * rad, sinTheta * sinPhi * rad, cosTheta * rad);
        float phase = TWOPI * distance;
        p.pol = mk(xcos(phase), xsin(phase), 0.f);
        rotateUz(p.pol, p.mom);
    } else if (type == 6) {
        float u0z, u1a, u, v, b, a;
        int turns = 0;
        do {
            u0z = zx + uniform(r) * (zy - zx);
            u1a = ax + uniform(r) * (ay - ax);
            u = 2.f * u0z - 1.f;
            v = 2.f * u1a - 1.f;
            b = u * u + v * v;
        } while (b > 1.f && ++turns < MAX_REJECT);
        a = 2.f * sqrtf(1.f - b);
        float rad = fabsf(radius);
        float flip = copysignf(1.f, radius);
        p.mom = mk(flip * a * u, flip * a * v, flip * (2.f * b - 1.f));
        p.pos = mk(a * u * rad, a * v * rad, (2.f * b - 1.f) * rad);
        float phase = TWOPI * distance;
        p.pol = mk(xcos(phase), xsin(phase), 0.f);
        rotateUz(p.pol, p.mom);
    } else if (type == 2) {
        p.mom = gmom;
        float frac = float(photon_id) / float(numphoton);
        float sfrac = 2.f * (frac - 0.5f);
        float rr = radius * sfrac;
        p.pos = mk(rr, 0.f, 0.f);
        rotateUz(p.pos, p.mom);
        p.pos = add(p.pos, gpos);
        p.pol = mk(0.f, -1.f, 0.f);
        rotateUz(p.pol, p.mom);
    } else if (type == 3) {
        p.mom = gmom;
        p.pos = mk(0.f, 0.f, 0.f);
        rotateUz(p.pos, p.mom);
        p.pos = add(p.pos, gpos);
        p.pol = mk(0.f, -1.f, 0.f);
        rotateUz(p.pol, p.mom);
    } else if (type == 4) {
        float f = float(photon_id) / float(numphoton);
        float frac = ax * (1.f - f) + ay * f;
        float phi = TWOPI * frac;
        float sinPhi = xsin(phi), cosPhi = xcos(phi);
        float rr = radius < 0.f ? -radius : radius;
        p.mom = mk(radius < 0.f ? -cosPhi : cosPhi, 0.f, radius < 0.f ? -sinPhi : sinPhi);
        p.pos = mk(rr * cosPhi, 0.f, rr * sinPhi);
        p.pos = add(p.pos, gpos);
        p.pol = mk(0.f, -1.f, 0.f);
        rotateUz(p.pol, p.mom);
    } else if (type == 5) {
        int side_size = int(numphoton / 4);
        int side = int(photon_id) / side_size;
        int side_index = int(photon_id) - side * side_size;
        float frac = float(side_index) / float(side_size);
        if (side == 0 || side == 1) {
            p.pos = mk(side == 0 ? ax : ay, 0.f, (1.f - frac) * zx + frac * zy);
            p.mom = mk(side == 0 ? 1.f : -1.f, 0.f, 0.f);
        } else if (side == 2 || side == 3) {
            p.pos = mk((1.f - frac) * ax + frac * ay, 0.f, side == 2 ? zx : zy);
            p.mom = mk(0.f, 0.f, side == 2 ? 1.f : -1.f);
        }
        p.pos = add(p.pos, gpos);
        p.pol = mk(0.f, -1.f, 0.f);
        rotateUz(p.pol, p.mom);
    } else if (type == 8) {
        float u0 = uniform(r), u1 = uniform(r), u2 = uniform(r);
        float phi = TWOPI * u0;
        float cosTheta = 1.f - 2.f * u1;
        float sinTheta = sqrtf(1.f - cosTheta * cosTheta);
        p.mom = mk(sinTheta * xcos(phi), sinTheta * xsin(phi), cosTheta);
        p.pos = gpos;
        float phase = TWOPI * u2;
        p.pol = mk(xcos(phase), xsin(phase), 0.f);
        rotateUz(p.pol, p.mom);
    }
    zero_flags(p);
    set_flag(p, TORCH);
}

// qcerenkov::generate (qudarap/qcerenkov.h:56-327)
OXG void gen_cerenkov(Ph& p, const float4* gs, Rng& r, const Tables& t) {
    float4 q0 = gs[0], q1 = gs[1], q2 = gs[2], q3 = gs[3], q4 = gs[4], q5 = gs[5];
    f3 dp = mk(q2.x, q2.y, q2.z);
    f3 p0 = normalize(dp);
    unsigned matline = __float_as_uint(q0.z);
    float BetaInverse = q4.x, Wmin = q4.y, Wmax = q4.z, maxSin2 = q5.x, mean1 = q5.y, mean2 = q5.z;
    float wavelength, cosTheta, sin2Theta, u_maxSin2;
    unsigned count = 0;
    do {
        float u0 = uniform(r);
        float w = Wmin + u0 * (Wmax - Wmin);
        wavelength = Wmin * Wmax / w;
        float4 props = boundary_lookup(t, wavelength, matline, 0u);
        float sampledRI = props.x;
        cosTheta = BetaInverse / sampledRI;
        sin2Theta = fmaxf(0.f, (1.f - cosTheta) * (1.f + cosTheta));
        float u1 = uniform(r);
        u_maxSin2 = u1 * maxSin2;
        count += 1;
    } while (u_maxSin2 > sin2Theta && count < 100);
    float sinTheta = sqrtf(sin2Theta);
    float u0 = uniform(r);
    float phi = 2.f * PIf * u0;
    float sinPhi = xsin(phi), cosPhi = xcos(phi);
    p.mom = mk(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
    rotateUz(p.mom, p0);
    p.pol = mk(cosTheta * cosPhi, cosTheta * sinPhi, -sinTheta);
    rotateUz(p.pol, p0);
    p.wavelength = wavelength;
    float fraction, delta, N, NumberOfPhotons;
    float MeanMax = fmaxf(mean1, mean2);
    float DeltaN = mean1 - mean2;
    float step_length = q2.w;
    int turns = 0;
    do {
        fraction = uniform(r);
        delta = fraction * step_length;
        NumberOfPhotons = mean1 - fraction * DeltaN;
        float u = uniform(r);
        N = u * MeanMax;
    } while (N > NumberOfPhotons && ++turns < MAX_REJECT);
    float preV = q3.w, postV = q5.w;
    float midVelocity = preV + fraction * (postV - preV) * 0.5f;
    p.time = q1.w + delta / midVelocity;
    p.pos = add(mk(q1.x, q1.y, q1.z), mul(fraction, dp));
    zero_flags(p);
    set_flag(p, CERENKOV);
}

// qscint::generate + mom_pol_wavelength (qudarap/qscint.h:55-145)
OXG void gen_scint(Ph& p, const float4* gs, Rng& r, const Tables& t) {
    float4 q1 = gs[1], q2 = gs[2], q3 = gs[3], q5 = gs[5];
    float u0 = uniform(r), u1 = uniform(r), u2 = uniform(r), u3 = uniform(r);
    float cost = 1.f - 2.f * u0;
    float sint = sqrtf((1.f - cost) * (1.f + cost));
    float phi = 2.f * PIf * u1;
    float sinp = xsin(phi), cosp = xcos(phi);
    p.mom = mk(sint * cosp, sint * sinp, cost);
    p.pol = mk(cost * cosp, cost * sinp, -sint);
    phi = 2.f * PIf * u2;
    sinp = xsin(phi); cosp = xcos(phi);
    p.pol = normalize(add(mul(cosp, p.pol), mul(sinp, cross(p.mom, p.pol))));
    p.wavelength = scint_wavelength(t, u3);
    float charge = q3.y;
    float fraction = charge == 0.f ? 1.f : uniform(r);
    p.pos = add(mk(q1.x, q1.y, q1.z), mul(fraction, mk(q2.x, q2.y, q2.z)));
    float u4 = uniform(r);
    float deltaTime = fraction * q2.w / q3.w - q5.x * xlog(u4);
    p.time = q1.w + deltaTime;
    zero_flags(p);
    set_flag(p, SCINTILLATION);
}

} // namespace oxd
