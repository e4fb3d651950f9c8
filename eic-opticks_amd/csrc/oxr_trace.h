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
//   flat[]  int2 (instance, local prim) per leaf when the geometry is flattened: one BVH
//           over every placed prim's world box, no BLAS level. trace<..., FLAT=true>
//           walks that tree; FLAT=false walks the two-level one.
#pragma once
#include "oxr_device.h"

namespace oxd {

struct BNode { float4 a, b, c; int4 ch; };
// A staged node is read through an LDS-typed pointer, so it loads with ds_read; a load
// through one pointer that may address either LDS or global memory compiles to a flat load,
// which takes the vector-memory path for both.
typedef float LdsF4 __attribute__((ext_vector_type(4)));
typedef int LdsI4 __attribute__((ext_vector_type(4)));
struct LdsNode { LdsF4 a, b, c; LdsI4 ch; };
typedef __attribute__((address_space(3))) const LdsNode LdsNodeL;
OXD BNode lds_node(const BNode* top, int i) {
    LdsNodeL* p = (LdsNodeL*)top + i;
    const LdsF4 a = p->a, b = p->b, c = p->c;
    const LdsI4 ch = p->ch;
    BNode n;
    n.a = make_float4(a.x, a.y, a.z, a.w);
    n.b = make_float4(b.x, b.y, b.z, b.w);
    n.c = make_float4(c.x, c.y, c.z, c.w);
    n.ch = make_int4(ch.x, ch.y, ch.z, ch.w);
    return n;
}
struct InstD { float4 c0, c1, c2; int4 info; };   // info: prim offset, ii_id, flags (bit 0 identity,
                                                  // bits 8-15 visibility mask), blas_root
// an instance takes part in a trace when its visibility mask shares a bit with the ray's
// (OptixInstance::visibilityMask against optixTrace's visibilityMask, SBT.cc:538-545)
OXD bool visible(const InstD& in, unsigned vizmask) { return ((unsigned(in.info.z) >> 8) & vizmask) != 0u; }
struct SolidD { int primOffset, numPrim, blasRoot, pad; };

struct Geom {
    const float4* node;
    const float4* plan;
    const float4* itra;
    const int2* prim;
    const BNode* tlas;
    const BNode* blas;
    const InstD* inst;
    const int2* flat;      // flattened leaves (instance, local prim), or null for two levels
    int tlas_root;
    int num_inst;
    unsigned vizmask;      // ray visibility mask against InstD flag bits 8-15
    // radial bounds, two float4 per node {centre, (r_in0, r_out0, r_in1, r_out1)} of the
    // top tree (instance tree or flattened tree) and of the per-solid trees
    const float4* rad_top;
    const float4* rad_blas;
};

constexpr int SENTINEL = int(0x80000000u);

struct Hit {
    float t;
    f3 n;          // world normal (not normalised)
    f3 lpos;       // object-frame hit position
    uint32_t ii_id, gpbd;
};

// Slab-test helpers. These only decide which boxes are entered, never a hit's value:
// the reciprocal is the hardware approximation (v_rcp_f32, 1 ulp) and the slab
// distances are fma(bound, inv, -origin*inv). The host widens every stored box by more
// than the resulting rounding for ray origins anywhere in the scene.
OXD f3 safe_inv(f3 d) {
    const float e = 1e-20f;
    return mk(__builtin_amdgcn_rcpf(fabsf(d.x) > e ? d.x : copysignf(e, d.x)),
              __builtin_amdgcn_rcpf(fabsf(d.y) > e ? d.y : copysignf(e, d.y)),
              __builtin_amdgcn_rcpf(fabsf(d.z) > e ? d.z : copysignf(e, d.z)));
}
OXD f3 neg_oi(f3 o, f3 inv) { return mk(-(o.x * inv.x), -(o.y * inv.y), -(o.z * inv.z)); }
// slab test of one child box (noi = -origin*inv)
OXD bool box_hit(float lx, float ly, float lz, float hx, float hy, float hz, f3 noi, f3 inv, float tbest, float& tn) {
    float tx0 = fmaf(lx, inv.x, noi.x), tx1 = fmaf(hx, inv.x, noi.x);
    float ty0 = fmaf(ly, inv.y, noi.y), ty1 = fmaf(hy, inv.y, noi.y);
    float tz0 = fmaf(lz, inv.z, noi.z), tz1 = fmaf(hz, inv.z, noi.z);
    float n = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
    float f = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    tn = n;
    return n <= f && f >= 0.f && n <= tbest;
}

// world -> object frame of one instance (rows c0..c2 hold the world-to-object 3x4)
OXD void to_object(const InstD& in, f3 pos, f3 mom, f3& o, f3& d) {
    o = mk(in.c0.x * pos.x + in.c0.y * pos.y + in.c0.z * pos.z + in.c0.w,
           in.c1.x * pos.x + in.c1.y * pos.y + in.c1.z * pos.z + in.c1.w,
           in.c2.x * pos.x + in.c2.y * pos.y + in.c2.z * pos.z + in.c2.w);
    d = mk(in.c0.x * mom.x + in.c0.y * mom.y + in.c0.z * mom.z,
           in.c1.x * mom.x + in.c1.y * mom.y + in.c1.z * mom.z,
           in.c2.x * mom.x + in.c2.y * mom.y + in.c2.z * mom.z);
}

// Closest hit in (tmin, tmax]; ties in t go to the smaller (instance, local prim) key.
// SHALLOW: every CSG tree of the geometry has height <= 2 (register-resident CSG stacks).
// RADIAL: besides its box, each child must pass its radial bound: its surfaces lie between
// r_in and r_out from the node's centre c, so along a ray from o (direction length |d|) they
// are no nearer than max(r_in - |o - c|, |o - c| - r_out) / |d|; a child whose bound lies
// beyond the best hit is skipped, and the bound orders the children and fills the distance
// stack. Concentric shells, whose boxes all hold the ray origin, are culled this way.
// stk points at this lane's LDS stack column (element k at stk[k*stride]). With TN,
// a second column of cap floats follows the stack (at stk + stride*cap) holding each
// pushed node's box entry distance, and a popped node whose entry lies beyond the best
// hit so far is dropped without loading it: the same cull the box test applies, so the
// hit found is the same.
// During traversal only (t, instance, local prim, object-frame normal) of the best hit
// stay live; its world normal, local position and identity words are formed afterwards
// from the instance record, with the same arithmetic as at instance entry.
template <bool FP64, bool COUNT, bool LIST, bool TN = false, bool FLAT = false, bool SHALLOW = false, bool RADIAL = false,
          bool SIMPLE = false>
OXD bool trace(const Geom& g, f3 pos, f3 mom, float tmin, float tmax, Hit& hit,
               int* stk, int stride, int cap, const BNode* top, int ntop,
               uint32_t& visits, uint32_t& tests, uint32_t& err, uint32_t& csgb, uint32_t& ivis) {
    float best = tmax;
    int best_inst = 0, best_item = 0;
    f3 bn = mk(0.f, 0.f, 0.f);
    bool found = false;
    int sp = 0;
    int cur = g.tlas_root;
    bool at_top = true;
    f3 o = pos, d = mom, inv = safe_inv(mom), noi = neg_oi(pos, inv);
    const float ilen_world = RADIAL ? __builtin_amdgcn_rsqf(dot(mom, mom)) : 1.f;
    float ilen = ilen_world;
    float* stkt = reinterpret_cast<float*>(stk + stride * cap);
    // pop the next entry worth visiting; false when the stack is empty
    auto pop = [&]() -> bool {
        while (sp > 0) {
            sp--;
            if (!TN || stkt[sp * stride] <= best) { cur = stk[sp * stride]; return true; }
        }
        return false;
    };
    int inst = 0;
    int primOffset = 0;
    for (;;) {
        if (cur >= 0) {
            BNode nd;
            if (at_top && cur < ntop) nd = lds_node(top, cur);
            else nd = at_top ? g.tlas[cur] : g.blas[cur];
            if (COUNT) visits++;
            float tn0, tn1;
            bool h0 = box_hit(nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, noi, inv, best, tn0);
            bool h1 = box_hit(nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, noi, inv, best, tn1);
            if (RADIAL) {
                const float4* rb = at_top ? g.rad_top : g.rad_blas;
                const float4 rc = rb[2 * cur], rr = rb[2 * cur + 1];
                const float dx = o.x - rc.x, dy = o.y - rc.y, dz = o.z - rc.z;
                const float dist = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
                const float lb0 = fmaxf(rr.x - dist, dist - rr.y) * ilen;
                const float lb1 = fmaxf(rr.z - dist, dist - rr.w) * ilen;
                h0 = h0 && lb0 <= best;
                h1 = h1 && lb1 <= best;
                tn0 = fmaxf(tn0, lb0);
                tn1 = fmaxf(tn1, lb1);
            }
            if (h0 && h1) {
                bool first0 = tn0 <= tn1;
                int nearc = first0 ? nd.ch.x : nd.ch.y;
                int farc = first0 ? nd.ch.y : nd.ch.x;
                if (sp < cap) {
                    stk[sp * stride] = farc;
                    if (TN) stkt[sp * stride] = first0 ? tn1 : tn0;
                    sp++;
                } else {
                    err |= 1u;
                }
                cur = nearc;
            } else if (h0) {
                cur = nd.ch.x;
            } else if (h1) {
                cur = nd.ch.y;
            } else if (!pop()) {
                break;
            }
            continue;
        }
        if (cur == SENTINEL) {
            at_top = true;
            o = pos; d = mom; inv = safe_inv(mom); noi = neg_oi(pos, inv);
            ilen = ilen_world;
            if (!pop()) break;
            continue;
        }
        int item = ~cur;
        if (FLAT) {
            // flattened leaf: one placed prim, tested in its instance's frame; the same
            // (instance, local prim) key and arithmetic as the two-level path
            const int2 fi = g.flat[item];
            const InstD in = g.inst[fi.x];
            if (COUNT) ivis++;
            if (!visible(in, g.vizmask)) {
                if (!pop()) break;
                continue;
            }
            f3 lo = pos, ld = mom;
            if ((in.info.z & 1) == 0) to_object(in, pos, mom, lo, ld);
            int nodeOffset = g.prim[in.info.x + fi.y].x;
            if (COUNT) {
                tests++;
                unsigned tc0 = __float_as_uint(g.node[4 * nodeOffset + 3].z);
                unsigned nn = tc0 >= CSG_LEAF ? 1u : __float_as_uint(g.node[4 * nodeOffset].x);
                for (unsigned k = 0; k < nn; k++) {
                    float4 q3 = g.node[4 * (nodeOffset + int(k)) + 3];
                    csgb += 64u;
                    if (__float_as_uint(q3.z) >= CSG_LEAF && (__float_as_uint(q3.w) & 0x7fffffffu) != 0u) csgb += 64u;
                }
            }
            Is is;
            if (intersect_prim<FP64, LIST, SHALLOW, SIMPLE>(is, g.node, nodeOffset, g.plan, g.itra, tmin, lo, ld)) {
                float t = is.w;
                bool closer = t >= tmin && (t < best || (found && t == best && (fi.x < best_inst || (fi.x == best_inst && fi.y < best_item))));
                if (closer) {
                    found = true;
                    best = t; best_inst = fi.x; best_item = fi.y;
                    bn = mk(is.x, is.y, is.z);
                }
            }
            if (!pop()) break;
            continue;
        }
        if (at_top) {
            const InstD in = g.inst[item];
            if (COUNT) ivis++;
            if (!visible(in, g.vizmask)) {
                if (!pop()) break;
                continue;
            }
            inst = item;
            primOffset = in.info.x;
            if ((in.info.z & 1) == 0) {
                to_object(in, pos, mom, o, d);
                inv = safe_inv(d);
                noi = neg_oi(o, inv);
                if (RADIAL) ilen = __builtin_amdgcn_rsqf(dot(d, d));
            }
            if (sp < cap) {
                stk[sp * stride] = SENTINEL;
                if (TN) stkt[sp * stride] = -INFINITY;
                sp++;
            } else {
                err |= 1u;
            }
            at_top = false;
            cur = in.info.w;
            continue;
        }
        // prim leaf: item is the local prim index within the instance's solid
        int nodeOffset = g.prim[primOffset + item].x;
        if (COUNT) {
            tests++;
            unsigned tc0 = __float_as_uint(g.node[4 * nodeOffset + 3].z);
            unsigned nn = tc0 >= CSG_LEAF ? 1u : __float_as_uint(g.node[4 * nodeOffset].x);
            for (unsigned k = 0; k < nn; k++) {
                float4 q3 = g.node[4 * (nodeOffset + int(k)) + 3];
                csgb += 64u;
                if (__float_as_uint(q3.z) >= CSG_LEAF && (__float_as_uint(q3.w) & 0x7fffffffu) != 0u) csgb += 64u;
            }
        }
        Is is;
        if (intersect_prim<FP64, LIST, SHALLOW, SIMPLE>(is, g.node, nodeOffset, g.plan, g.itra, tmin, o, d)) {
            // a reported distance outside the ray interval is ignored, as by
            // optixReportIntersection (CSGOptiX/CSGOptiX7.cu:912): a convex polyhedron
            // wholly behind the ray reports valid with the zeroed isect
            float t = is.w;
            bool closer = t >= tmin && (t < best || (found && t == best && (inst < best_inst || (inst == best_inst && item < best_item))));
            if (closer) {
                found = true;
                best = t; best_inst = inst; best_item = item;
                bn = mk(is.x, is.y, is.z);
            }
        }
        if (!pop()) break;
    }
    if (!found) {
        hit.t = 0.f; hit.n = mk(0.f, 0.f, 0.f); hit.lpos = mk(0.f, 0.f, 0.f);
        hit.ii_id = 0xffffffffu; hit.gpbd = 0xffffffffu;
        return false;
    }
    const InstD in = g.inst[best_inst];
    f3 bo = pos, bd = mom;
    bool ident = (in.info.z & 1) != 0;
    if (!ident) to_object(in, pos, mom, bo, bd);
    hit.t = best;
    hit.lpos = mk(bo.x + best * bd.x, bo.y + best * bd.y, bo.z + best * bd.z);
    hit.ii_id = uint32_t(in.info.y);
    hit.gpbd = uint32_t(g.prim[in.info.x + best_item].y);
    if (!ident) {
        hit.n = mk(in.c0.x * bn.x + in.c1.x * bn.y + in.c2.x * bn.z,
                   in.c0.y * bn.x + in.c1.y * bn.y + in.c2.y * bn.z,
                   in.c0.z * bn.x + in.c1.z * bn.y + in.c2.z * bn.z);
    } else {
        hit.n = bn;
    }
    return true;
}

} // namespace oxd
