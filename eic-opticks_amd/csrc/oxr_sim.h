// oxr_sim.h -- the per-photon pieces and the two simulate kernels (plain and persistent
// form), templated on root precision (FP64), counting (COUNT), debug outputs (DBG), list-node
// support (LIST) and the traversal mode TM (bit 0: stack entries carry box entry distances;
// bit 1: the flattened BVH; bit 2: every CSG tree has height <= 2, so the CSG evaluation
// stacks live in registers, production kernels without list nodes only; bit 3: nodes also
// cull by their radial bounds, production and counting kernels). k_generate (oxr_kernels.hip) uses the per-photon pieces; the
// kernels are instantiated by the four oxr_sim_*.hip units, one per (FP64, LIST) pair, which
// compile in parallel.
#pragma once
#include "oxr_kernels.h"


// Minimum waves per SIMD the simulate kernels are compiled for (VGPR budget 512/N per
// lane). The latency-bound traversal wants occupancy more than registers: measured on
// MI355X (crystal-calorimeter event), 4 waves beat 3 and 5 for both forms; with the
// simple-leaf kernels 5 waves lost on every event (instanced CSG 11.4 against 8.8 ms), and
// with the CSG arrays in LDS 3 waves lost too (instanced CSG 9.0 against 7.9 ms).
#ifndef OXR_PERSIST_WAVES
#define OXR_PERSIST_WAVES 4
#endif
#ifndef OXR_PLAIN_WAVES
#define OXR_PLAIN_WAVES 4
#endif

namespace oxd {

// ------------------------------------------------------------------ seeding
// genstep g with offsets[g] <= i < offsets[g+1] (offsets non-decreasing, i < offsets[num_gs])
OXD int genstep_of(const long long* __restrict__ offsets, int num_gs, long long i) {
    int lo = 0, hi = num_gs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (offsets[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------------ per-photon pieces
struct Lane {
    Ph p;
    Rng r;
    unsigned long long sh0, sh1, sb0, sb1;   // seqhis / seqbnd
    float lc, lf;                            // lposcost, lposfphi of the last trace (photonlite)
    long long idx;
    int bounce;
};

template <bool DBG>
OXD void point(const Params& P, Lane& L) {
    if (!DBG) return;
    if (P.record && L.bounce < P.max_record) store_photon(P.record + 4 * (L.idx * P.max_record + L.bounce), L.p);
    if (P.rec && L.bounce < P.max_rec) P.rec[L.idx * P.max_rec + L.bounce] = rec_pack(L.p, P.extent_domain, P.time_domain);
    if (P.seq) {
        unsigned slot = unsigned(L.bounce);
        if (slot < 32) {
            unsigned shift = 4u * (slot & 15u);
            unsigned long long his = (unsigned long long)(unsigned(__ffs(int(ph_flag(L.p)))) & 0xfu) << shift;
            unsigned long long bnd = (unsigned long long)(ph_boundary(L.p) & 0xfu) << shift;
            if (slot < 16) { L.sh0 |= his; L.sb0 |= bnd; } else { L.sh1 |= his; L.sb1 |= bnd; }
        }
    }
}

// the LDS copy of the node and itra arrays and the prim table (TM bit 5): after the P.ntop
// staged top nodes
// and the traversal stacks (blockDim.x lanes x P.stack_cap entries of 4 B, 8 B with
// distances), the layout simulate_smem sizes
OXD const float4* csg_copy(const Params& P) {
    extern __shared__ int4 smem4[];
    return reinterpret_cast<const float4*>(smem4) + 4 * P.ntop + (blockDim.x * unsigned(P.stack_cap) * (P.tn_stack ? 2u : 1u)) / 4u;
}

// returns false for padding indices of a batched step
template <bool DBG>
OXD bool begin_photon(const Params& P, Lane& L, long long idx) {
    L.idx = idx;
    const float4* gsb = P.gs;
    unsigned long long pidx;
    int gsi;
    int evt = P.event_index;
    if (P.segs) {
        int lo = 0, hi = P.nseg - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (P.segs[mid].start <= idx) lo = mid; else hi = mid - 1;
        }
        const Seg sg = P.segs[lo];
        long long li = idx - sg.start;
        if (li >= sg.count) return false;
        pidx = sg.photon0 + (unsigned long long)li;
        gsb = sg.gs;
        gsi = genstep_of(sg.off, sg.num_gs, (long long)pidx);   // served events carry no seed array
        evt = sg.event_index;
    } else {
        pidx = P.slot_offset + (unsigned long long)idx;
        gsi = P.seed[idx];
    }
    const float4* gs = gsb + 6 * gsi;
    rng_init(L.r, P.rng_seed, P.rng_offset, P.skip, (unsigned long long)evt, pidx);
    if (DBG && P.pre) { L.r.pre = P.pre + (size_t)idx * P.npre; L.r.npre = P.npre; }
    uint32_t gencode = __float_as_uint(gs[0].x);
    // the photon starts zeroed (sctx ctx = {}, CSGOptiX7.cu:432): a generator that leaves a
    // field unset (T_RECTANGLE beyond its fourth side sets no position or direction,
    // sysrap/storch.h:460-512) leaves it zero
    L.p = Ph{};
    if (gencode == 6) gen_torch(L.p, gs, pidx, L.r);
    else if (gencode == 15 || gencode == 18) gen_cerenkov(L.p, gs, L.r, P.t);
    else if (gencode == 16 || gencode == 5) gen_scint(L.p, gs, L.r, P.t);
    else if (gencode == 19) { load_photon(L.p, P.input + 4 * pidx); set_flag(L.p, TORCH); }
    else { zero_flags(L.p); }
    set_index(L.p, pidx);
    L.sh0 = L.sh1 = L.sb0 = L.sb1 = 0ull;
    if (DBG) { L.lc = 0.f; L.lf = 0.f; }
    L.bounce = 0;
    point<DBG>(P, L);
    return true;
}

// trace<refine> (CSGOptiX7.cu:110-216): with PropagateRefine, when 0.99 of the distance
// found lies beyond the refine distance, the ray is traced again from that point and the
// 0.99 part added back, so a far surface is intersected from close by. A miss reports
// distance 1, as the miss program writes it (CSGOptiX7.cu:666-670), so h.t is the
// reference's prd distance in every case. One trace call site serves both passes.
// CSGL: the node and itra arrays and the prim table are read from the LDS copy stage_csg made.
template <bool FP64, bool COUNT, bool LIST, bool TN, bool FLAT, bool SHALLOW, bool RADIAL, bool SIMPLE = false,
          bool CSGL = false>
OXD bool trace_refined(const Params& P, f3 pos, f3 mom, float tmin, Hit& h, int* stk, int stride,
                       const BNode* top, int ntop, uint32_t& visits, uint32_t& tests, uint32_t& err,
                       uint32_t& csgb, uint32_t& ivis) {
    Geom g = P.g;
    if constexpr (CSGL) {
        g.node = csg_copy(P);
        g.itra = g.node + P.csg_node4;
        g.prim = reinterpret_cast<const int2*>(g.itra + P.csg_itra4);
    }
    f3 o = pos;
    float t_approx = 0.f;
    bool refined = false, found;
    for (;;) {
        found = trace<FP64, COUNT, LIST, TN, FLAT, SHALLOW, RADIAL, SIMPLE>(g, o, mom, tmin, P.tmax, h, stk, stride, P.stack_cap,
                                                                   top, ntop, visits, tests, err, csgb, ivis);
        if (!found) h.t = 1.f;
        if (!P.refine || refined) break;
        t_approx = 0.99f * h.t;
        if (!(t_approx > P.refine_distance)) break;
        o = mk(pos.x + t_approx * mom.x, pos.y + t_approx * mom.y, pos.z + t_approx * mom.z);
        refined = true;
    }
    if (refined) h.t = h.t + t_approx;
    return found;
}

// one loop iteration of CSGOptiX7.cu:450-482; returns false when the photon is done
template <bool FP64, bool COUNT, bool DBG, bool LIST, int TM>
OXD bool step_photon(const Params& P, Lane& L, int* stk, int stride, const BNode* top, int ntop,
                     uint32_t& visits, uint32_t& tests, uint32_t& err, uint32_t& csgb, uint32_t& ivis) {
    if (!(L.bounce < P.max_bounce && L.p.time < P.max_time)) return false;
    float tmin = (L.p.obf & P.eps0_mask) ? P.tmin0 : P.tmin;
    Hit h;
    bool found = trace_refined<FP64, COUNT, LIST, (TM & 1) != 0, (TM & 2) != 0, (TM & 4) != 0, (TM & 8) != 0, (TM & 16) != 0,
                               (TM & 32) != 0>(
        P, L.p.pos, L.p.mom, tmin, h, stk, stride, top, ntop, visits, tests, err, csgb, ivis);
    if (!found) {
        if (DBG) { L.lc = 0.f; L.lf = 0.f; }   // __miss__ms sets lposcost = lposfphi = 0
        return false;
    }
    // normalize(n) is n * (1 / sqrt(dot(n, n))): when that dot rounds to exactly 1 it
    // returns n itself, so an already unit normal (boxes, polyhedra) skips the arithmetic
    const float nn = dot(h.n, h.n);
    f3 normal = nn == 1.f ? h.n : normalize(h.n);
    if (DBG && (P.lite || P.prd)) {
        L.lc = lpos_cost(h.lpos);
        L.lf = (xatan2(h.lpos.y, h.lpos.x) + PIf) / (2.0f * PIf);   // normalize_fphi, sysrap/scuda.h:653
    }
    // sctx::trace (sysrap/sctx.h:170-173): the trace result of this bounce, normal normalised
    // (CSGOptiX7.cu:466-474), in the quad2 layout of sysrap/squad.h:95-112
    if (DBG && P.prd && L.bounce < P.max_prd) {
        float4* q = P.prd + 2 * (L.idx * P.max_prd + L.bounce);
        q[0] = make_float4(normal.x, normal.y, normal.z, h.t);
        q[1] = make_float4(L.lc, L.lf, __uint_as_float(h.ii_id), __uint_as_float(h.gpbd));
    }
    int command = propagate(L.p, h, normal, L.r, P.t, P.aligned);
    L.bounce++;
    point<DBG>(P, L);
    return command != BREAK;
}

// production launches: pick up a photon that k_generate wrote (false for padding)
OXD bool resume_photon(const Params& P, Lane& L, long long idx) {
    const ulonglong2 st = P.rstate[idx];
    if (st.x == ~0ull) return false;
    L.idx = idx;
    load_photon(L.p, P.photon + 4 * idx);
    rng_resume(L.r, P.rng_seed, st.x, st.y);
    L.bounce = 0;
    return true;
}

template <bool DBG>
OXD bool start_photon(const Params& P, Lane& L, long long idx) {
    if (DBG) return begin_photon<true>(P, L, idx);
    if (P.gen_inplace) return begin_photon<false>(P, L, idx);
    return resume_photon(P, L, idx);
}

template <bool DBG>
OXD void end_photon(const Params& P, const Lane& L) {
    if (DBG && P.seq) {
        ulonglong2* s = P.seq + 2 * L.idx;
        s[0] = make_ulonglong2(L.sh0, L.sh1);
        s[1] = make_ulonglong2(L.sb0, L.sb1);
    }
    store_photon(P.photon + 4 * L.idx, L.p);
    // all bits of the hit mask must be set (sphoton_selector, sysrap/sphoton.h:1219-1225)
    if (P.hitflag) P.hitflag[L.idx] = (L.p.flagmask & P.hitmask) == P.hitmask ? 1 : 0;
    if (DBG && P.lite) P.lite[L.idx] = lite_pack(L.p, L.lc, L.lf);
    if (DBG && P.bounce_out) P.bounce_out[L.idx] = uint32_t(L.bounce);
    if (DBG && P.overflow_out) P.overflow_out[L.idx] = L.r.overflow;
}

OXD void stage_top(const Params& P, BNode* top, int ntop) {
    for (int i = threadIdx.x; i < ntop; i += blockDim.x) top[i] = P.g.tlas[i];
}

// TM bit 5 (32): the node and itra arrays and the prim table are copied into LDS after the
// top nodes and the traversal stacks, and the trace reads them there. The copy's address derives from the
// LDS array alone, so its reads compile to ds_read.
template <int TM>
OXD void stage_csg(const Params& P) {
    if constexpr ((TM & 32) != 0) {
        float4* l = const_cast<float4*>(csg_copy(P));
        for (int i = threadIdx.x; i < P.csg_node4; i += blockDim.x) l[i] = P.g.node[i];
        for (int i = threadIdx.x; i < P.csg_itra4; i += blockDim.x) l[P.csg_node4 + i] = P.g.itra[i];
        int2* lp = reinterpret_cast<int2*>(l + P.csg_node4 + P.csg_itra4);
        for (int i = threadIdx.x; i < P.csg_prim; i += blockDim.x) lp[i] = P.g.prim[i];
    }
}

OXD unsigned long long wave_sum(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

OXD void flush_counters(const Params& P, unsigned long long steps, unsigned long long visits,
                        unsigned long long tests, uint32_t err, bool count,
                        unsigned long long csgb = 0, unsigned long long ivis = 0) {
    steps = wave_sum(steps);
    if (count) { visits = wave_sum(visits); tests = wave_sum(tests); csgb = wave_sum(csgb); ivis = wave_sum(ivis); }
    uint64_t anyerr = __ballot(err != 0u);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(P.counters + 0, steps);
        if (count) {
            atomicAdd(P.counters + 2, visits); atomicAdd(P.counters + 3, tests);
            atomicAdd(P.counters + 5, csgb); atomicAdd(P.counters + 6, ivis);
        }
        if (anyerr) atomicAdd(P.counters + 4, 1ull);
    }
}


// ------------------------------------------------------------------ plain form
template <bool FP64, bool COUNT, bool DBG, bool LIST, int TM>
__global__ void __launch_bounds__(256, OXR_PLAIN_WAVES) k_simulate(Params P) {
    extern __shared__ int4 smem4[];
    BNode* top = reinterpret_cast<BNode*>(smem4);
    int* stk = reinterpret_cast<int*>(top + P.ntop) + threadIdx.x;
    stage_top(P, top, P.ntop);
    stage_csg<TM>(P);
    __syncthreads();
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long steps = 0, visits = 0, tests = 0, csgb = 0, ivis = 0;
    uint32_t err = 0;
    if (idx < P.num_photon) {
        Lane L;
        if (start_photon<DBG>(P, L, idx)) {
            uint32_t v = 0, t = 0, cb = 0, iv = 0;
            while (step_photon<FP64, COUNT, DBG, LIST, TM>(P, L, stk, blockDim.x, top, P.ntop, v, t, err, cb, iv)) {}
            end_photon<DBG>(P, L);
            steps = (unsigned long long)L.bounce; visits = v; tests = t; csgb = cb; ivis = iv;
        } else {
            P.hitflag[idx] = 0;
        }
    }
    flush_counters(P, steps, visits, tests, err, COUNT, csgb, ivis);
}

// ------------------------------------------------------------------ persistent form
// Lanes hold one photon each and run one bounce step per loop turn. When at least
// refill_threshold lanes of a wave are idle (or all are), the wave claims that many
// photon indices with one atomicAdd on counters[1] and the idle lanes generate them;
// with P.perm the k-th claim is photon P.perm[k] (a sorted order), else photon k.
// Every wave leaves the loop once the counter has passed num_photon and its lanes are idle.
template <bool FP64, bool DBG, bool LIST, int TM>
__global__ void __launch_bounds__(256, OXR_PERSIST_WAVES) k_simulate_persist(Params P) {
    extern __shared__ int4 smem4[];
    BNode* top = reinterpret_cast<BNode*>(smem4);
    int* stk = reinterpret_cast<int*>(top + P.ntop) + threadIdx.x;
    stage_top(P, top, P.ntop);
    stage_csg<TM>(P);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    Lane L;
    bool active = false;
    bool drained = false;
    unsigned long long steps = 0;
    uint32_t v = 0, t = 0, err = 0, cb = 0, iv = 0;
    for (;;) {
        uint64_t idle = __ballot(!active);
        int nidle = __popcll(idle);
        if (!drained && nidle > 0 && (nidle >= P.refill_threshold || nidle == 64)) {
            int leader = __ffsll((long long)idle) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(P.counters + 1, (unsigned long long)nidle);
            base = __shfl(base, leader, 64);
            if (base + (unsigned long long)nidle >= (unsigned long long)P.num_photon) drained = true;
            if (!active) {
                unsigned rank = __builtin_amdgcn_mbcnt_hi(uint32_t(idle >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(idle), 0u));
                unsigned long long my = base + rank;
                if (my < (unsigned long long)P.num_photon) {
                    const long long pi = P.perm ? (long long)P.perm[my] : (long long)my;
                    if (start_photon<DBG>(P, L, pi)) active = true;
                    else P.hitflag[pi] = 0;
                }
            }
        }
        if (__ballot(active) == 0ull) {
            if (drained) break;
            continue;
        }
        if (active) {
            if (!step_photon<FP64, false, DBG, LIST, TM>(P, L, stk, blockDim.x, top, P.ntop, v, t, err, cb, iv)) {
                end_photon<DBG>(P, L);
                steps += (unsigned long long)L.bounce;
                active = false;
            }
        }
    }
    flush_counters(P, steps, 0, 0, err, false);
}

} // namespace oxd

namespace oxl {

// kind: SIM_PROD (production), SIM_DEBUG (seqhis, records, precooked streams), SIM_COUNT
// (visit counting, generating in place like the debug variants). tm: traversal mode bits
// (oxd kernels above); debug launches use the node-index stack without radial culling,
// counting launches the production traversal without the register CSG stacks (which
// change no visit), so they count the visits the production launch makes.
template <bool FP64, bool LIST>
hipError_t launch_simulate_plain(hipStream_t s, const oxd::Params& P, dim3 g, dim3 b, size_t sm, int kind, int tm) {
#define OXR_PLAIN(C, D, T) hipLaunchKernelGGL((oxd::k_simulate<FP64, C, D, LIST, T>), g, b, sm, s, P)
    if (kind == SIM_COUNT) {
        switch (tm & 11) {
            case 0: OXR_PLAIN(true, true, 0); break;
            case 1: OXR_PLAIN(true, true, 1); break;
            case 2: OXR_PLAIN(true, true, 2); break;
            case 3: OXR_PLAIN(true, true, 3); break;
            case 8: OXR_PLAIN(true, true, 8); break;
            case 9: OXR_PLAIN(true, true, 9); break;
            case 10: OXR_PLAIN(true, true, 10); break;
            default: OXR_PLAIN(true, true, 11); break;
        }
    } else if (kind == SIM_DEBUG) {
        if (tm & 2) OXR_PLAIN(false, true, 2);
        else OXR_PLAIN(false, true, 0);
    } else {
        switch (tm & 11) {
            case 0: if ((tm & 4) && !LIST) OXR_PLAIN(false, false, 4 * !LIST); else OXR_PLAIN(false, false, 0); break;
            case 1: if ((tm & 4) && !LIST) OXR_PLAIN(false, false, 1 + 4 * !LIST); else OXR_PLAIN(false, false, 1); break;
            case 2: if ((tm & 4) && !LIST) OXR_PLAIN(false, false, 2 + 4 * !LIST); else OXR_PLAIN(false, false, 2); break;
            case 3: if ((tm & 4) && !LIST) OXR_PLAIN(false, false, 3 + 4 * !LIST); else OXR_PLAIN(false, false, 3); break;
            case 8: if ((tm & 4) && !LIST) OXR_PLAIN(false, false, 8 + 4 * !LIST); else OXR_PLAIN(false, false, 8); break;
            case 9: if ((tm & 4) && !LIST) OXR_PLAIN(false, false, 9 + 4 * !LIST); else OXR_PLAIN(false, false, 9); break;
            case 10: if ((tm & 4) && !LIST) OXR_PLAIN(false, false, 10 + 4 * !LIST); else OXR_PLAIN(false, false, 10); break;
            default: if ((tm & 4) && !LIST) OXR_PLAIN(false, false, 11 + 4 * !LIST); else OXR_PLAIN(false, false, 11); break;
        }
    }
#undef OXR_PLAIN
    return hipGetLastError();
}

// the persistent kernel instance launch_simulate_persist runs for (debug, tm)
template <bool FP64, bool LIST>
const void* persist_kernel(bool debug, int tm) {
#define OXR_K(D, T) reinterpret_cast<const void*>(&oxd::k_simulate_persist<FP64, D, LIST, T>)
    if (debug) return (tm & 2) ? OXR_K(true, 2) : OXR_K(true, 0);
    const bool sh = (tm & 4) && !LIST;
    switch (tm & 11) {
        case 0: return sh ? OXR_K(false, 4 * !LIST) : OXR_K(false, 0);
        case 1: return sh ? OXR_K(false, 1 + 4 * !LIST) : OXR_K(false, 1);
        case 2: return sh ? OXR_K(false, 2 + 4 * !LIST) : OXR_K(false, 2);
        case 3: return sh ? OXR_K(false, 3 + 4 * !LIST) : OXR_K(false, 3);
        case 8: return sh ? OXR_K(false, 8 + 4 * !LIST) : OXR_K(false, 8);
        case 9: return sh ? OXR_K(false, 9 + 4 * !LIST) : OXR_K(false, 9);
        case 10: return sh ? OXR_K(false, 10 + 4 * !LIST) : OXR_K(false, 10);
        default: return sh ? OXR_K(false, 11 + 4 * !LIST) : OXR_K(false, 11);
    }
#undef OXR_K
}

template <bool FP64, bool LIST>
hipError_t launch_simulate_persist(hipStream_t s, const oxd::Params& P, dim3 g, dim3 b, size_t sm, bool debug, int tm) {
#define OXR_PERSIST(D, T) hipLaunchKernelGGL((oxd::k_simulate_persist<FP64, D, LIST, T>), g, b, sm, s, P)
    if (debug) {
        if (tm & 2) OXR_PERSIST(true, 2);
        else OXR_PERSIST(true, 0);
    } else {
        switch (tm & 11) {
            case 0: if ((tm & 4) && !LIST) OXR_PERSIST(false, 4 * !LIST); else OXR_PERSIST(false, 0); break;
            case 1: if ((tm & 4) && !LIST) OXR_PERSIST(false, 1 + 4 * !LIST); else OXR_PERSIST(false, 1); break;
            case 2: if ((tm & 4) && !LIST) OXR_PERSIST(false, 2 + 4 * !LIST); else OXR_PERSIST(false, 2); break;
            case 3: if ((tm & 4) && !LIST) OXR_PERSIST(false, 3 + 4 * !LIST); else OXR_PERSIST(false, 3); break;
            case 8: if ((tm & 4) && !LIST) OXR_PERSIST(false, 8 + 4 * !LIST); else OXR_PERSIST(false, 8); break;
            case 9: if ((tm & 4) && !LIST) OXR_PERSIST(false, 9 + 4 * !LIST); else OXR_PERSIST(false, 9); break;
            case 10: if ((tm & 4) && !LIST) OXR_PERSIST(false, 10 + 4 * !LIST); else OXR_PERSIST(false, 10); break;
            default: if ((tm & 4) && !LIST) OXR_PERSIST(false, 11 + 4 * !LIST); else OXR_PERSIST(false, 11); break;
        }
    }
#undef OXR_PERSIST
    return hipGetLastError();
}

template <int BASE>
hipError_t launch_simulate_simple(hipStream_t s, const oxd::Params& P, dim3 g, dim3 b, size_t sm, bool persist, int tm) {
#define OXR_S(T)                                                                                          \
    do {                                                                                                  \
        if (persist) hipLaunchKernelGGL((oxd::k_simulate_persist<false, false, false, BASE + T>), g, b, sm, s, P); \
        else hipLaunchKernelGGL((oxd::k_simulate<false, false, false, false, BASE + T>), g, b, sm, s, P);  \
    } while (0)
    switch (tm & 15) {
        case 0: OXR_S(0); break;   case 1: OXR_S(1); break;   case 2: OXR_S(2); break;   case 3: OXR_S(3); break;
        case 4: OXR_S(4); break;   case 5: OXR_S(5); break;   case 6: OXR_S(6); break;   case 7: OXR_S(7); break;
        case 8: OXR_S(8); break;   case 9: OXR_S(9); break;   case 10: OXR_S(10); break; case 11: OXR_S(11); break;
        case 12: OXR_S(12); break; case 13: OXR_S(13); break; case 14: OXR_S(14); break; default: OXR_S(15); break;
    }
#undef OXR_S
    return hipGetLastError();
}

template <int BASE>
const void* persist_kernel_simple(int tm) {
#define OXR_K(T) return reinterpret_cast<const void*>(&oxd::k_simulate_persist<false, false, false, BASE + T>)
    switch (tm & 15) {
        case 0: OXR_K(0);   case 1: OXR_K(1);   case 2: OXR_K(2);   case 3: OXR_K(3);
        case 4: OXR_K(4);   case 5: OXR_K(5);   case 6: OXR_K(6);   case 7: OXR_K(7);
        case 8: OXR_K(8);   case 9: OXR_K(9);   case 10: OXR_K(10); case 11: OXR_K(11);
        case 12: OXR_K(12); case 13: OXR_K(13); case 14: OXR_K(14); default: OXR_K(15);
    }
#undef OXR_K
}

} // namespace oxl
