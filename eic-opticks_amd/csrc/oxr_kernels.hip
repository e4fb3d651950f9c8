// oxr_kernels.hip -- gfx950 kernels of the engine and their host-side launchers.
//
//   k_fill_seed        per-photon genstep index (QEvt.cu:181-237 reduce + iexpand)
//   k_simulate         plain form: one lane per photon, CSGOptiX7.cu:405-503 shape
//   k_simulate_persist persistent form: a fixed grid whose lanes refill from a device
//                      photon counter between bounce steps (wave-aggregated atomics)
//   k_count/k_scan/k_scatter  ordered hit compaction (SU::count_if/copy_if, SU.cu:49-166)
//   k_trace_rays, k_intersect_prim  geometry queries for tests and the simtrace-like API
#include "oxr_kernels.h"

// Minimum waves per SIMD the simulate kernels are compiled for (VGPR budget 512/N per
// lane). The latency-bound traversal wants occupancy more than registers: measured on
// MI355X (crystal-calorimeter event), 4 waves beat 3 and 5 for both forms.
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

__global__ void k_fill_seed(const long long* __restrict__ offsets, int num_gs, int* __restrict__ seed, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    seed[i] = genstep_of(offsets, num_gs, i);
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
    if (gencode == 6) gen_torch(L.p, gs, pidx, L.r);
    else if (gencode == 15 || gencode == 18) gen_cerenkov(L.p, gs, L.r, P.t);
    else if (gencode == 16 || gencode == 5) gen_scint(L.p, gs, L.r, P.t);
    else if (gencode == 19) { load_photon(L.p, P.input + 4 * pidx); set_flag(L.p, TORCH); }
    else { L.p = Ph{}; zero_flags(L.p); }
    set_index(L.p, pidx);
    L.sh0 = L.sh1 = L.sb0 = L.sb1 = 0ull;
    if (DBG) { L.lc = 0.f; L.lf = 0.f; }
    L.bounce = 0;
    point<DBG>(P, L);
    return true;
}

// one loop iteration of CSGOptiX7.cu:450-482; returns false when the photon is done
template <bool FP64, bool COUNT, bool DBG>
OXD bool step_photon(const Params& P, Lane& L, int* stk, int stride, const BNode* top, int ntop,
                     uint32_t& visits, uint32_t& tests, uint32_t& err, uint32_t& csgb, uint32_t& ivis) {
    if (!(L.bounce < P.max_bounce && L.p.time < P.max_time)) return false;
    float tmin = (L.p.obf & P.eps0_mask) ? P.tmin0 : P.tmin;
    Hit h;
    bool found = trace<FP64, COUNT>(P.g, L.p.pos, L.p.mom, tmin, P.tmax, h, stk, stride, P.stack_cap, top, ntop,
                                    visits, tests, err, csgb, ivis);
    if (!found) {
        if (DBG) { L.lc = 0.f; L.lf = 0.f; }   // __miss__ms sets lposcost = lposfphi = 0
        return false;
    }
    f3 normal = normalize(h.n);
    float lposcost = h.lpos.z / sqrtf(dot(h.lpos, h.lpos));
    if (DBG && P.lite) {
        L.lc = lposcost;
        L.lf = (xatan2(h.lpos.y, h.lpos.x) + PIf) / (2.0f * PIf);   // normalize_fphi, sysrap/scuda.h:653
    }
    int command = propagate(L.p, h, normal, lposcost, L.r, P.t, P.aligned);
    L.bounce++;
    point<DBG>(P, L);
    return command != BREAK;
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
template <bool FP64, bool COUNT, bool DBG>
__global__ void __launch_bounds__(256, OXR_PLAIN_WAVES) k_simulate(Params P) {
    extern __shared__ int4 smem4[];
    BNode* top = reinterpret_cast<BNode*>(smem4);
    int* stk = reinterpret_cast<int*>(top + P.ntop) + threadIdx.x;
    stage_top(P, top, P.ntop);
    __syncthreads();
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long steps = 0, visits = 0, tests = 0, csgb = 0, ivis = 0;
    uint32_t err = 0;
    if (idx < P.num_photon) {
        Lane L;
        if (begin_photon<DBG>(P, L, idx)) {
            uint32_t v = 0, t = 0, cb = 0, iv = 0;
            while (step_photon<FP64, COUNT, DBG>(P, L, stk, blockDim.x, top, P.ntop, v, t, err, cb, iv)) {}
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
// photon indices with one atomicAdd on counters[1] and the idle lanes generate them.
// Every wave leaves the loop once the counter has passed num_photon and its lanes are idle.
template <bool FP64, bool DBG>
__global__ void __launch_bounds__(256, OXR_PERSIST_WAVES) k_simulate_persist(Params P) {
    extern __shared__ int4 smem4[];
    BNode* top = reinterpret_cast<BNode*>(smem4);
    int* stk = reinterpret_cast<int*>(top + P.ntop) + threadIdx.x;
    stage_top(P, top, P.ntop);
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
                    if (begin_photon<DBG>(P, L, (long long)my)) active = true;
                    else P.hitflag[my] = 0;
                }
            }
        }
        if (__ballot(active) == 0ull) {
            if (drained) break;
            continue;
        }
        if (active) {
            if (!step_photon<FP64, false, DBG>(P, L, stk, blockDim.x, top, P.ntop, v, t, err, cb, iv)) {
                end_photon<DBG>(P, L);
                steps += (unsigned long long)L.bounce;
                active = false;
            }
        }
    }
    flush_counters(P, steps, 0, 0, err, false);
}

// ------------------------------------------------------------------ hit compaction
constexpr int CHUNK = 4096;   // photons per compaction block (256 lanes x 16)

__global__ void __launch_bounds__(256) k_count(const uint8_t* __restrict__ flag, long long n, int* __restrict__ counts) {
    long long base = (long long)blockIdx.x * CHUNK;
    int c = 0;
    for (int k = threadIdx.x; k < CHUNK; k += 256) {
        long long i = base + k;
        if (i < n && flag[i]) c++;
    }
    c = (int)wave_sum((unsigned long long)c);
    __shared__ int ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// single workgroup exclusive scan of the per-block counts; total to out[nb]
__global__ void __launch_bounds__(1024) k_scan(const int* __restrict__ counts, int nb, long long* __restrict__ out) {
    __shared__ long long part[1024];
    long long carry = 0;
    for (int base = 0; base < nb; base += 1024) {
        int i = base + threadIdx.x;
        long long v = i < nb ? counts[i] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            long long add = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < nb) out[i] = carry + part[threadIdx.x] - v;
        long long tot = part[1023];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) out[nb] = carry;
}

// W float4 per item: 4 for sphoton, 1 for sphotonlite
template <int W>
__global__ void __launch_bounds__(256) k_scatter(const uint8_t* __restrict__ flag, const float4* __restrict__ photon,
                                                 long long n, const long long* __restrict__ offs,
                                                 float4* __restrict__ hit, long long hit_cap) {
    __shared__ int ws[4];
    long long base = (long long)blockIdx.x * CHUNK;
    long long o = offs[blockIdx.x];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = 0; k < CHUNK; k += 256) {
        long long i = base + k + threadIdx.x;
        bool f = i < n && flag[i];
        uint64_t m = __ballot(f);
        unsigned rank = __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
        if (lane == 0) ws[w] = __popcll(m);
        __syncthreads();
        int before = 0;
        for (int q = 0; q < w; q++) before += ws[q];
        int tot = ws[0] + ws[1] + ws[2] + ws[3];
        if (f) {
            long long dst = o + before + rank;
            if (dst < hit_cap) {
                const float4* s = photon + W * i;
                float4* d = hit + W * dst;
#pragma unroll
                for (int q = 0; q < W; q++) d[q] = s[q];
            }
        }
        o += tot;
        __syncthreads();
    }
}

// ------------------------------------------------------------------ queries
template <bool FP64>
__global__ void __launch_bounds__(256) k_trace_rays(Params P, const float4* __restrict__ rays, long long n, float4* __restrict__ prd) {
    extern __shared__ int4 smem4[];
    BNode* top = reinterpret_cast<BNode*>(smem4);
    int* stk = reinterpret_cast<int*>(top + P.ntop) + threadIdx.x;
    stage_top(P, top, P.ntop);
    __syncthreads();
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = rays[2 * i], b = rays[2 * i + 1];
    Hit h;
    uint32_t v = 0, t = 0, err = 0, cb = 0, iv = 0;
    bool found = trace<FP64, false>(P.g, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, b.w, h, stk, blockDim.x,
                                    P.stack_cap, top, P.ntop, v, t, err, cb, iv);
    float lc = 0.f, lf = 0.f;
    if (found) {
        lc = h.lpos.z / sqrtf(dot(h.lpos, h.lpos));
        lf = (xatan2(h.lpos.y, h.lpos.x) + PIf) / (2.0f * PIf);
    }
    prd[2 * i] = make_float4(h.n.x, h.n.y, h.n.z, h.t);
    prd[2 * i + 1] = make_float4(lc, lf, __uint_as_float(h.ii_id), __uint_as_float(h.gpbd));
}

template <bool FP64>
__global__ void __launch_bounds__(256) k_intersect_prim(Geom g, int root, const float4* __restrict__ rays, long long n,
                                                         float4* __restrict__ isect, int* __restrict__ valid) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = rays[2 * i], b = rays[2 * i + 1];
    Is is;
    bool v = intersect_prim<FP64>(is, g.node, root, g.plan, g.itra, a.w, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z));
    isect[i] = make_float4(is.x, is.y, is.z, is.w);
    valid[i] = v ? 1 : 0;
}

} // namespace oxd

// ====================================================================== launchers
namespace oxl {
using namespace oxd;

hipError_t fill_seed(hipStream_t s, const long long* offsets, int num_gs, int* seed, long long n) {
    if (n <= 0) return hipSuccess;
    long long nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_fill_seed, dim3((unsigned)nb), dim3(256), 0, s, offsets, num_gs, seed, n);
    return hipGetLastError();
}

size_t simulate_smem(const Params& P, int block) {
    return size_t(P.ntop) * sizeof(BNode) + size_t(block) * P.stack_cap * sizeof(int);
}

// debug outputs (seqhis, records, precooked streams) select the DBG kernel variants
static bool wants_debug(const Params& P) {
    return P.seq || P.record || P.pre || P.bounce_out || P.overflow_out || P.lite;
}

hipError_t simulate_plain(hipStream_t s, const Params& P, int block, bool fp64, bool count) {
    if (P.num_photon <= 0) return hipSuccess;
    long long nb = (P.num_photon + block - 1) / block;
    size_t sm = simulate_smem(P, block);
    dim3 g((unsigned)nb), b(block);
    if (count) {
        if (fp64) hipLaunchKernelGGL((k_simulate<true, true, true>), g, b, sm, s, P);
        else hipLaunchKernelGGL((k_simulate<false, true, true>), g, b, sm, s, P);
    } else if (wants_debug(P)) {
        if (fp64) hipLaunchKernelGGL((k_simulate<true, false, true>), g, b, sm, s, P);
        else hipLaunchKernelGGL((k_simulate<false, false, true>), g, b, sm, s, P);
    } else {
        if (fp64) hipLaunchKernelGGL((k_simulate<true, false, false>), g, b, sm, s, P);
        else hipLaunchKernelGGL((k_simulate<false, false, false>), g, b, sm, s, P);
    }
    return hipGetLastError();
}

hipError_t simulate_persist(hipStream_t s, const Params& P, int block, int grid, bool fp64) {
    if (P.num_photon <= 0) return hipSuccess;
    size_t sm = simulate_smem(P, block);
    bool dbg = wants_debug(P);
    if (fp64) {
        if (dbg) hipLaunchKernelGGL((k_simulate_persist<true, true>), dim3(grid), dim3(block), sm, s, P);
        else hipLaunchKernelGGL((k_simulate_persist<true, false>), dim3(grid), dim3(block), sm, s, P);
    } else {
        if (dbg) hipLaunchKernelGGL((k_simulate_persist<false, true>), dim3(grid), dim3(block), sm, s, P);
        else hipLaunchKernelGGL((k_simulate_persist<false, false>), dim3(grid), dim3(block), sm, s, P);
    }
    return hipGetLastError();
}

int compact_blocks(long long n) { return int((n + CHUNK - 1) / CHUNK); }

hipError_t compact(hipStream_t s, const uint8_t* flag, const float4* photon, long long n, int* counts,
                   long long* offs, float4* hit, long long hit_cap) {
    if (n <= 0) return hipSuccess;
    int nb = compact_blocks(n);
    hipLaunchKernelGGL(k_count, dim3(nb), dim3(256), 0, s, flag, n, counts);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, counts, nb, offs);
    hipLaunchKernelGGL(k_scatter<4>, dim3(nb), dim3(256), 0, s, flag, photon, n, offs, hit, hit_cap);
    return hipGetLastError();
}

hipError_t scatter_lite(hipStream_t s, const uint8_t* flag, const uint4* lite, long long n, const long long* offs,
                        uint4* hitlite, long long hit_cap) {
    if (n <= 0) return hipSuccess;
    int nb = compact_blocks(n);
    hipLaunchKernelGGL(k_scatter<1>, dim3(nb), dim3(256), 0, s, flag, reinterpret_cast<const float4*>(lite), n, offs,
                       reinterpret_cast<float4*>(hitlite), hit_cap);
    return hipGetLastError();
}

hipError_t trace_rays(hipStream_t s, const Params& P, const float4* rays, long long n, float4* prd, bool fp64) {
    if (n <= 0) return hipSuccess;
    long long nb = (n + 255) / 256;
    size_t sm = simulate_smem(P, 256);
    if (fp64) hipLaunchKernelGGL((k_trace_rays<true>), dim3((unsigned)nb), dim3(256), sm, s, P, rays, n, prd);
    else hipLaunchKernelGGL((k_trace_rays<false>), dim3((unsigned)nb), dim3(256), sm, s, P, rays, n, prd);
    return hipGetLastError();
}

hipError_t intersect_prim(hipStream_t s, const Geom& g, int root, const float4* rays, long long n, float4* isect,
                          int* valid, bool fp64) {
    if (n <= 0) return hipSuccess;
    long long nb = (n + 255) / 256;
    if (fp64) hipLaunchKernelGGL((k_intersect_prim<true>), dim3((unsigned)nb), dim3(256), 0, s, g, root, rays, n, isect, valid);
    else hipLaunchKernelGGL((k_intersect_prim<false>), dim3((unsigned)nb), dim3(256), 0, s, g, root, rays, n, isect, valid);
    return hipGetLastError();
}

} // namespace oxl
