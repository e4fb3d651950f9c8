// oxr_kernels.h -- kernel parameter block and launcher declarations shared by the
// kernels (oxr_kernels.hip) and the host engine (oxr_engine.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "oxr_physics.h"

namespace oxd {

// One event's contiguous photon range inside a batched serving step. Step indices
// [start, start+count) belong to this event; [start+count, next start) is padding so
// that every segment begins on a compaction chunk.
struct Seg {
    long long start, count;
    unsigned long long photon0;   // photon index within the event of the first photon
    const float4* gs;
    const long long* off;         // the event's genstep photon offsets (num_gs + 1)
    int num_gs;
    int event_index;
};

// Everything one simulate launch reads or writes (the role of CSGOptiX/Params.h:29-122
// plus the device sevent, sysrap/sevent.h:105-233).
struct Params {
    Geom g;
    Tables t;
    const float4* gs;          // quad6 gensteps
    const int* seed;           // genstep index per photon
    const float4* input;       // input photons (INPUT_PHOTON gensteps)
    const float* pre;          // precooked uniforms (num_photon, npre) or null
    int npre;
    float4* photon;            // sphoton out
    ulonglong2* seq;           // sseq out (2 x ulonglong2 per photon) or null
    float4* record;            // records or null
    int max_record;
    uint4* rec;                // srec out (16 B per (photon, bounce) slot) or null
    int max_rec;
    float4* prd;               // quad2 prd out (2 float4 per slot) or null
    int max_prd;
    float extent_domain, time_domain;   // rec compression domains (mm, ns)
    uint4* lite;               // sphotonlite out (16 B per photon) or null
    ulonglong2* rstate;        // generate -> propagate hand-over: {next uniform, photon index};
                               // x = ~0 marks a padding slot of a batched step
    uint8_t* hitflag;          // 1 if (flagmask & hitmask) == hitmask
    uint32_t* bounce_out;      // optional per-photon loop count
    uint32_t* overflow_out;    // optional precooked-stream wrap count
    unsigned long long* counters;   // [0] steps [1] queue [2] node visits [3] prim tests [4] errors
    long long num_photon;
    unsigned long long slot_offset;
    int event_index;
    int max_bounce;
    float tmin, tmin0;
    unsigned eps0_mask;
    float tmax, max_time;
    unsigned hitmask;
    unsigned long long rng_seed, rng_offset, skip;
    int aligned;
    int ntop;                  // top-level BVH nodes staged in LDS
    int stack_cap;             // per-lane traversal stack entries in LDS
    int tn_stack;              // stack entries carry their box entry distance (8 B each)
    int shallow_csg;           // every CSG tree has height <= 2 and no prim holds a list node
    int radial;                // nodes also cull by their radial bounds (tuning bvh_radial)
    int simple_leaves;         // every leaf is a sphere, zsphere, cylinder or box (no list nodes)
    int csg_node4, csg_itra4;  // > 0: the simple-leaf kernels copy the node and itra arrays (this many
    int csg_prim;              // float4) and the prim table (this many int2) into LDS after the
                               // traversal stacks and read them there
    const unsigned* perm;      // persistent form: the order photons are claimed in, or null (index order)
    int gen_inplace;           // production kernels generate each photon where they claim it
    int refine;                // PropagateRefine: trace again from 0.99 of a far hit
    float refine_distance;     // PropagateRefineDistance (mm)
    int refill_threshold;      // persistent form: idle lanes per wave before a refill
    const Seg* segs;           // batched serving step, or null for a single event
    int nseg;
};

} // namespace oxd

namespace oxl {
hipError_t fill_seed(hipStream_t s, const long long* offsets, int num_gs, int* seed, long long n);
size_t simulate_smem(const oxd::Params& P, int block);
// Production launches (no debug outputs, no counting) run generate() first: the
// propagate kernels then resume the photons it wrote.
bool needs_generate(const oxd::Params& P, bool count);
hipError_t generate(hipStream_t s, const oxd::Params& P);
hipError_t simulate_plain(hipStream_t s, const oxd::Params& P, int block, bool fp64, bool count, bool list);
hipError_t simulate_persist(hipStream_t s, const oxd::Params& P, int block, int grid, bool fp64, bool list);
// The simulate kernel variants of one (FP64 roots, list support) pair, defined in oxr_sim.h
// and instantiated in oxr_sim_{f32,f64}[_list].hip. kind: SIM_PROD, SIM_DEBUG or SIM_COUNT;
// tm: traversal mode, bit 0 distance-carrying stack, bit 1 flattened BVH, bit 2 register CSG
// stacks, bit 3 radial culling.
enum { SIM_PROD = 0, SIM_DEBUG = 1, SIM_COUNT = 2 };
template <bool FP64, bool LIST>
hipError_t launch_simulate_plain(hipStream_t s, const oxd::Params& P, dim3 g, dim3 b, size_t sm, int kind, int tm);
template <bool FP64, bool LIST>
hipError_t launch_simulate_persist(hipStream_t s, const oxd::Params& P, dim3 g, dim3 b, size_t sm, bool debug, int tm);
#define OXR_SIM_EXTERN(F, L)                                                                                        \
    extern template hipError_t launch_simulate_plain<F, L>(hipStream_t, const oxd::Params&, dim3, dim3, size_t, int, int); \
    extern template hipError_t launch_simulate_persist<F, L>(hipStream_t, const oxd::Params&, dim3, dim3, size_t, bool, int);
OXR_SIM_EXTERN(false, false)
OXR_SIM_EXTERN(false, true)
OXR_SIM_EXTERN(true, false)
OXR_SIM_EXTERN(true, true)
#undef OXR_SIM_EXTERN
// looks up the device address of the no-alloc scratch array for `device` (oxr_noalloc.hip)
hipError_t noalloc_setup(int device);
// production kernels of geometries whose leaves are all spheres, zspheres, cylinders or
// boxes, with single-precision roots and no list nodes: the traversal mode tm (bits 0-3)
// plus BASE, 16 (the simple-leaf bit; oxr_sim_f32_simple.hip) or 48 (also bit 5: node and
// itra arrays and the prim table read from their LDS copy, P.csg_node4 > 0;
// oxr_sim_f32_simple_lds.hip)
template <int BASE>
hipError_t launch_simulate_simple(hipStream_t s, const oxd::Params& P, dim3 g, dim3 b, size_t sm, bool persist, int tm);
template <int BASE>
const void* persist_kernel_simple(int tm);
#define OXR_SIMPLE_EXTERN(B)                                                                                                 \
    extern template hipError_t launch_simulate_simple<B>(hipStream_t, const oxd::Params&, dim3, dim3, size_t, bool, int); \
    extern template const void* persist_kernel_simple<B>(int);
OXR_SIMPLE_EXTERN(16)
OXR_SIMPLE_EXTERN(48)
#undef OXR_SIMPLE_EXTERN
// workgroups of block lanes per CU that can be resident at once for the persistent
// kernel simulate_persist would launch for P (-1 if the runtime cannot tell)
int persist_occupancy(const oxd::Params& P, int block, bool fp64, bool list);
int compact_blocks(long long n);
// true when P launches the production kernels (no seqhis, records, rec, prd, photonlite
// or precooked streams)
bool production_kernels(const oxd::Params& P);
// The claim order of a production persistent launch, photons sorted by the key
// (genstep index << 16) | Morton code of the direction's octahedral cell (256 x 256), read
// from the generated photons; sorted on key bits [0, key_bits). key and val hold 2 x
// num_photon entries each, tmp *tmp_bytes; with tmp null only *tmp_bytes is set. *perm
// receives the sorted photon indices (inside val).
hipError_t order_photons(hipStream_t s, const oxd::Params& P, int key_bits, unsigned long long* key, unsigned* val,
                         void* tmp, size_t* tmp_bytes, const unsigned** perm);
hipError_t compact(hipStream_t s, const uint8_t* flag, const float4* photon, long long n, int* counts,
                   long long* offs, float4* hit, long long hit_cap);
// compact() for the serving form: the chunk offsets (nb+1) and the 8 launch counters are
// also written to host_offs/host_ctr, and hit rows below hcap to hhit, all pinned host
// memory; rows from hcap on go to dhit (device, dcap rows) at the same row index
hipError_t compact_host(hipStream_t s, const uint8_t* flag, const float4* photon, long long n, int* counts,
                        long long* offs, long long* host_offs, const unsigned long long* ctr,
                        unsigned long long* host_ctr, float4* hhit, long long hcap, float4* dhit, long long dcap);
// hit selection of the sphotonlite array with the offsets a compact() of the same flags left
hipError_t scatter_lite(hipStream_t s, const uint8_t* flag, const uint4* lite, long long n, const long long* offs,
                        uint4* hitlite, long long hit_cap);
// merge of sphoton (lite=false) or sphotonlite items by (identity, time bucket); see oxr_merge.hip
hipError_t merge(hipStream_t s, bool lite, const void* in, long long n, unsigned select_mask, bool select,
                 float time_window, void* out, long long* num_out);
hipError_t trace_rays(hipStream_t s, const oxd::Params& P, const float4* rays, long long n, float4* prd, bool fp64,
                      bool list);
hipError_t simtrace(hipStream_t s, const oxd::Params& P, const float4* gs, const long long* offsets, int num_gs,
                    long long n, float4* out, bool fp64, bool list);
hipError_t intersect_prim(hipStream_t s, const oxd::Geom& g, int root, const float4* rays, long long n,
                          float4* isect, int* valid, bool fp64, bool list);
} // namespace oxl
