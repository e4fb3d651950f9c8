// oxr_trace.h -- two-level BVH closest-hit traversal, the engine's replacement for the
// OptiX IAS/GAS traversal behind optixTrace (CSGOptiX/CSGOptiX7.cu:110-216) together
// with __intersection__is, __closesthit__ch and __miss__ms (CSGOptiX7.cu:655-940).
//
// Layout in HBM:
//   tlas[]  BNode 64 B: both child boxes + child refs; BFS order so the top levels are
//           the first nodes and a workgroup can stage them in LDS.
//   blas[]  BNode 64 B per-solid trees over prim AABBs, concatenated.
//   inst[]  InstD 64 B: world-to-object columns + (solid, iindex|identity, flags, blas root).
//   prim[]  int2 (nodeOffset, globalPrimIdx<<16|boundary).
// A child ref >= 0 is an internal node, < 0 is a leaf item ~ref (one instance per TLAS
// leaf, one prim per BLAS leaf). The traversal stack lives in LDS, one column per lane.
#pragma once
#include "oxr_device.h"

namespace oxd {

struct BNode { float4 a, b, c; int4 ch; };
struct InstD { float4 c0, c1, c2; int4 info; };   // info: solid, ii_id, flags(bit0 identity), blas_root
struct SolidD { int primOffset, numPrim, blasRoot, pad; };

struct Geom {
    const float4* node;
    const float4* plan;
    const float4* itra;
    const int2* prim;
    const BNode* tlas;
    const BNode* blas;
    const InstD* inst;
    int tlas_root;
    int num_inst;
};

constexpr int SENTINEL = int(0x80000000u);

struct Hit {
    float t;
    f3 n;          // world normal (not normalised)
    f3 lpos;       // object-frame hit position
    uint32_t ii_id, gpbd;
};

OXD f3 safe_inv(f3 d) {
    const float e = 1e-20f;
    return mk(1.f / (fabsf(d.x) > e ? d.x : copysignf(e, d.x)),
              1.f / (fabsf(d.y) > e ? d.y : copysignf(e, d.y)),
              1.f / (fabsf(d.z) > e ? d.z : copysignf(e, d.z)));
}

// slab test of one child box; boxes are pre-expanded on the host so this is conservative
OXD bool box_hit(float lx, float ly, float lz, float hx, float hy, float hz, f3 o, f3 inv, float tbest, float& tn) {
    float tx0 = (lx - o.x) * inv.x, tx1 = (hx - o.x) * inv.x;
    float ty0 = (ly - o.y) * inv.y, ty1 = (hy - o.y) * inv.y;
    float tz0 = (lz - o.z) * inv.z, tz1 = (hz - o.z) * inv.z;
    float n = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
    float f = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    tn = n;
    return n <= f && f >= 0.f && n <= tbest;
}

// Closest hit in (tmin, tmax]; ties in t go to the smaller (instance, local prim) key.
// stk points at this lane's LDS stack column (element k at stk[k*stride]).
template <bool FP64, bool COUNT>
OXD bool trace(const Geom& g, f3 pos, f3 mom, float tmin, float tmax, Hit& hit,
               int* stk, int stride, int cap, const BNode* top, int ntop,
               uint32_t& visits, uint32_t& tests, uint32_t& err, uint32_t& csgb, uint32_t& ivis) {
    float best = tmax;
    uint64_t best_key = ~0ull;
    bool found = false;
    int sp = 0;
    int cur = g.tlas_root;
    bool at_top = true;
    f3 o = pos, d = mom, inv = safe_inv(mom);
    int inst = 0;
    float4 c0 = make_float4(1.f, 0.f, 0.f, 0.f), c1 = make_float4(0.f, 1.f, 0.f, 0.f), c2 = make_float4(0.f, 0.f, 1.f, 0.f);
    bool ident = true;
    uint32_t ii_id = 0;
    int primOffset = 0;
    for (;;) {
        if (cur >= 0) {
            BNode nd;
            if (at_top) nd = cur < ntop ? top[cur] : g.tlas[cur];
            else nd = g.blas[cur];
            if (COUNT) visits++;
            float tn0, tn1;
            bool h0 = box_hit(nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, o, inv, best, tn0);
            bool h1 = box_hit(nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, o, inv, best, tn1);
            if (h0 && h1) {
                bool first0 = tn0 <= tn1;
                int nearc = first0 ? nd.ch.x : nd.ch.y;
                int farc = first0 ? nd.ch.y : nd.ch.x;
                if (sp < cap) { stk[sp * stride] = farc; sp++; } else { err |= 1u; }
                cur = nearc;
            } else if (h0) {
                cur = nd.ch.x;
            } else if (h1) {
                cur = nd.ch.y;
            } else {
                if (sp == 0) break;
                sp--; cur = stk[sp * stride];
            }
            continue;
        }
        if (cur == SENTINEL) {
            at_top = true;
            o = pos; d = mom; inv = safe_inv(mom);
            if (sp == 0) break;
            sp--; cur = stk[sp * stride];
            continue;
        }
        int item = ~cur;
        if (at_top) {
            InstD in = g.inst[item];
            if (COUNT) ivis++;
            inst = item;
            c0 = in.c0; c1 = in.c1; c2 = in.c2;
            ident = (in.info.z & 1) != 0;
            ii_id = uint32_t(in.info.y);
            primOffset = in.info.x;
            if (!ident) {
                o = mk(c0.x * pos.x + c0.y * pos.y + c0.z * pos.z + c0.w,
                       c1.x * pos.x + c1.y * pos.y + c1.z * pos.z + c1.w,
                       c2.x * pos.x + c2.y * pos.y + c2.z * pos.z + c2.w);
                d = mk(c0.x * mom.x + c0.y * mom.y + c0.z * mom.z,
                       c1.x * mom.x + c1.y * mom.y + c1.z * mom.z,
                       c2.x * mom.x + c2.y * mom.y + c2.z * mom.z);
                inv = safe_inv(d);
            }
            if (sp < cap) { stk[sp * stride] = SENTINEL; sp++; } else { err |= 1u; }
            at_top = false;
            cur = in.info.w;
            continue;
        }
        // prim leaf: item is the local prim index within the instance's solid
        int2 pr = g.prim[primOffset + item];
        if (COUNT) {
            tests++;
            unsigned tc0 = __float_as_uint(g.node[4 * pr.x + 3].z);
            unsigned nn = tc0 >= CSG_LEAF ? 1u : __float_as_uint(g.node[4 * pr.x].x);
            for (unsigned k = 0; k < nn; k++) {
                float4 q3 = g.node[4 * (pr.x + int(k)) + 3];
                csgb += 64u;
                if (__float_as_uint(q3.z) >= CSG_LEAF && (__float_as_uint(q3.w) & 0x7fffffffu) != 0u) csgb += 64u;
            }
        }
        Is is;
        if (intersect_prim<FP64>(is, g.node, pr.x, g.plan, g.itra, tmin, o, d)) {
            float t = is.w;
            uint64_t key = (uint64_t(uint32_t(inst)) << 32) | uint64_t(uint32_t(item));
            if (t < best || (found && t == best && key < best_key)) {
                found = true;
                best = t; best_key = key;
                hit.t = t;
                hit.lpos = mk(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
                hit.gpbd = uint32_t(pr.y);
                hit.ii_id = ii_id;
                if (!ident) {
                    hit.n = mk(c0.x * is.x + c1.x * is.y + c2.x * is.z,
                               c0.y * is.x + c1.y * is.y + c2.y * is.z,
                               c0.z * is.x + c1.z * is.y + c2.z * is.z);
                } else {
                    hit.n = mk(is.x, is.y, is.z);
                }
            }
        }
        if (sp == 0) break;
        sp--; cur = stk[sp * stride];
    }
    if (!found) {
        hit.t = 0.f; hit.n = mk(0.f, 0.f, 0.f); hit.lpos = mk(0.f, 0.f, 0.f);
        hit.ii_id = 0xffffffffu; hit.gpbd = 0xffffffffu;
    }
    return found;
}

} // namespace oxd
