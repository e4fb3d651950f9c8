/*
 * oxr.h -- C ABI of the MI355X-native optical-photon propagation engine.
 *
 * The engine replaces the GPU simulate path of eic-opticks:
 *   CSGOptiX (OptiX IAS/GAS + CSGOptiX7.cu simulate)    -> own two-level BVH + HIP kernels
 *   QSim / QEvt (genstep upload, seeding, launch, gather) -> oxr_event_* / oxr_simulate_launch
 *   CSGOptiXService::simulate(gs, eventID) -> hits       -> oxr_simulate / oxr_server_*
 *
 * Conventions
 *   - Every entry point returns 0 on success and a negative OXR_E* code on failure;
 *     the message is available from oxr_last_error() (thread-local). Nothing throws
 *     across this ABI (the reference throws CUDA_Exception, sysrap/CUDA_CHECK.h:8-45).
 *   - Plain pointers and sizes only. Host pointers unless the name says "device".
 *   - Array layouts are byte-identical to the reference's persisted .npy arrays:
 *       CSGFoundry  solid (ns,3,4)i32  prim (np,4,4)f32  node (nn,4,4)f32
 *                   plan (nl,4)f32     itra (nt,4,4)f32  inst (ni,4,4)f32
 *                   (CSG/CSGFoundry.h:456-462, CSG/CSGFoundry.cc:2768-2800)
 *       SSim        bnd (nb,4,2,nw,4)f32  optical (nb*4,4)u32  icdf (1|3,4096,1)f32
 *                   (qudarap/QBnd.cc:134-161, sysrap/sstandard.h:311,447, qudarap/QScint.cc:84-109)
 *       SEvt        genstep quad6 (n,6,4)f32 (sysrap/sevent.h:145-193)
 *                   photon/hit sphoton (n,4,4)f32 (sysrap/sphoton.h:171-192)
 *                   seq sseq (n,2,2)u64 = seqhis[2], seqbnd[2] (sysrap/sseq.h:54-65)
 *                   record (n,max_record,4,4)f32 (sysrap/sctx.h:134)
 *                   rec srec (n,max_rec), 16 B each: short4 post, char4 polw, uchar4 flag
 *                       (sysrap/srec.h:39-43, sevent::add_rec sysrap/sevent.h:614-622)
 *                   prd quad2 (n,max_prd,2,4)f32 (sysrap/sctx.h:170-173, squad.h:95-112)
 */
#ifndef OXR_H
#define OXR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OXR_VERSION_MAJOR 0
#define OXR_VERSION_MINOR 1

/* ---- status codes ---------------------------------------------------- */
enum {
    OXR_OK            =  0,
    OXR_EINVAL        = -1,   /* bad argument / inconsistent array shapes        */
    OXR_ENOMEM        = -2,   /* device or host allocation failed                 */
    OXR_EHIP          = -3,   /* HIP runtime error (message has hipGetErrorString)*/
    OXR_ECAPACITY     = -4,   /* caller buffer too small (count still returned)   */
    OXR_ESTATE        = -5,   /* call out of order (e.g. launch before gensteps)  */
    OXR_ETIMEOUT      = -6,   /* server wait timed out                            */
    OXR_ENOTFOUND     = -7    /* unknown request id / missing CSGFoundry file     */
};

/* ---- engine forms: plain per-launch, persistent lane-refill ----------- */
enum {
    OXR_FORM_PLAIN      = 0,  /* one lane per photon, CSGOptiX7.cu:405-503 shape    */
    OXR_FORM_PERSISTENT = 1   /* persistent grid, lanes refill from a device queue  */
};

/* ---- random-number consumption schedules: DEBUG_TAG vs PRODUCTION ----- */
enum {
    OXR_RNG_PRODUCTION = 0,   /* qsim.h built with PRODUCTION: no alignment burns     */
    OXR_RNG_ALIGNED    = 1    /* qsim.h built !PRODUCTION + DEBUG_TAG (reference's
                                 shipped kernel, CSGOptiX/CMakeLists.txt:54-56):
                                 +2 u per propagate_to_boundary (qsim.h:730-733),
                                 +1 per propagate_at_boundary, +4 more on reflect
                                 (qsim.h:1088-1090,1188-1202), +1 per
                                 propagate_at_surface (qsim.h:1687-1692)           */
};

/* ---- what oxr_event_get can gather ------------------------------------ */
enum {
    OXR_GET_PHOTON = 0,   /* sphoton[num_photon]                 QEvt::gatherPhoton  */
    OXR_GET_HIT    = 1,   /* sphoton[num_hit]                    QEvt::gatherHit     */
    OXR_GET_SEQ    = 2,   /* sseq[num_photon] (needs seq_enabled) QEvt::gatherSeq    */
    OXR_GET_RECORD = 3,   /* sphoton[num_photon*max_record]      QEvt::gatherRecord  */
    OXR_GET_SEED   = 4,   /* int32[num_photon] genstep index     QEvt seed buffer    */
    OXR_GET_GENSTEP= 5,   /* quad6[num_genstep]                                      */
    OXR_GET_PHOTONLITE = 6, /* sphotonlite[num_photon] (needs photonlite)
                               QEvt::gatherPhotonLite, sysrap/sphotonlite.h:47-52     */
    OXR_GET_HITLITE    = 7, /* sphotonlite[num_hit]: the photonlite items of the hits,
                               QEvt::gatherHitLite (qudarap/QEvt.cc:985-1019)         */
    OXR_GET_REC        = 8, /* srec[num_photon*max_rec] (needs max_rec > 0): the
                               domain-compressed step points of sctx::point
                               (sysrap/sctx.h:134-140, sevent::add_rec)               */
    OXR_GET_PRD        = 9  /* quad2[num_photon*max_prd] (needs max_prd > 0): the
                               trace result of each bounce with its normal normalised,
                               sctx::trace (sysrap/sctx.h:170-173); slots of bounces
                               that found no boundary stay zero                       */
};

/* ---- inputs ----------------------------------------------------------- */

/* CSGFoundry arrays (replaces CSGFoundry::upload, CSG/CSGFoundry.cc:3377-3404,
   and the OptiX GAS/IAS builds of CSGOptiX/SBT.cc:247-553). Instance i uses
   solid inst[i].q1.w (gas_idx) and reports identity inst[i].q2.w
   (sensor_identifier+1, 0 = not a sensor) as sqat4.h:345-372. */
typedef struct oxr_geometry {
    const int32_t* solid;  int64_t num_solid;  /* (ns,3,4) i32, CSGSolid 48 B       */
    const float*   prim;   int64_t num_prim;   /* (np,4,4) f32, CSGPrim 64 B        */
    const float*   node;   int64_t num_node;   /* (nn,4,4) f32, CSGNode 64 B        */
    const float*   plan;   int64_t num_plan;   /* (nl,4)   f32                       */
    const float*   itra;   int64_t num_itra;   /* (nt,4,4) f32 inverse CSG transforms*/
    const float*   inst;   int64_t num_inst;   /* (ni,4,4) f32 instance transforms   */
} oxr_geometry;

/* Physics tables (replaces QSim::UploadComponents, qudarap/QSim.cc:118-250). */
typedef struct oxr_physics {
    const float*    bnd;      /* (num_bnd,4,2,num_wl,4) f32                          */
    int32_t         num_bnd;
    int32_t         num_wl;   /* reference domain: 761 (60..820 nm, sysrap/sdomain.h) */
    float           wl_min;   /* nm of column 0 (60)                                   */
    float           wl_step;  /* nm per column (1)                                     */
    const uint32_t* optical;  /* (num_bnd*4,4) u32                                     */
    const float*    icdf;     /* (icdf_rows,icdf_width) f32 scintillation ICDF, or NULL*/
    int32_t         icdf_rows;   /* 1 or 3                                             */
    int32_t         icdf_width;  /* 4096                                               */
    int32_t         hd_factor;   /* 0, 10 or 20 (qudarap/qscint.h:150-228)             */
    float           reemission_scint_time; /* reserved, ignored: re-emission adds no delay
                                              (qsim.h:804-825)                           */
} oxr_physics;

/* Stepping and event controls (CSGOptiX/Params.h:29-122, sysrap/SEventConfig.cc). */
typedef struct oxr_config {
    int32_t  max_bounce;      /* SEventConfig MaxBounce, default 31                   */
    float    tmin;            /* PropagateEpsilon, default 0.05 mm                    */
    float    tmin0;           /* PropagateEpsilon0, default 0.05 mm                   */
    uint32_t eps0_mask;       /* PropagateEpsilon0Mask, default TO|CK|SI|SC|RE        */
    float    tmax;            /* default 1e6 mm                                       */
    float    max_time;        /* ns, default 1e27                                     */
    uint32_t hitmask;         /* default SURFACE_DETECT                               */
    uint64_t rng_seed;        /* Philox key                                           */
    uint64_t rng_offset;      /* Philox counter offset (in uniforms)                  */
    uint64_t skipahead_event; /* uniforms skipped per event index, default 100000     */
    int32_t  rng_schedule;    /* OXR_RNG_PRODUCTION | OXR_RNG_ALIGNED                 */
    int32_t  fp64_roots;      /* 1: robust quadratic roots in FP64 (sphere, zsphere,
                                 cylinder, cone), csg_robust_quadratic_roots.h:95-134 */
    int32_t  seq_enabled;     /* 1: write sseq seqhis/seqbnd per photon               */
    int32_t  max_record;      /* >0: write per-bounce sphoton records                 */
    int64_t  max_photon;      /* capacity of one launch slice (OPTICKS_MAX_SLOT); 0, the
                                 default, sizes it from the device memory as
                                 SEventConfig::HeuristicMaxSlot (SEventConfig.cc:1897-1919):
                                 87% of it over the config's bytes per photon, rounded
                                 down to a million; oxr_engine_get_config returns it   */
    int32_t  photonlite;      /* 1: also write sphotonlite (16 B) per photon, the last
                                 intersect's lposcost/lposfphi packed (SEventConfig
                                 ModeLite, CSGOptiX7.cu:493-499); default 0           */
    int32_t  max_rec;         /* >0: write per-bounce srec (SEventConfig MaxRec)      */
    int32_t  max_prd;         /* >0: write per-bounce quad2 prd (SEventConfig MaxPrd) */
    float    extent_domain;   /* rec position domain, mm, centred on the origin
                                 (SEventConfig MaxExtentDomain, default 1000)         */
    float    time_domain;     /* rec time domain, ns (MaxTimeDomain, default 10)      */
    uint32_t propagate_refine;/* 1: a trace whose distance d has 0.99 d beyond
                                 refine_distance is traced again from 0.99 d along the
                                 ray and 0.99 d added back (SEventConfig PropagateRefine,
                                 default 0; trace<true>, CSGOptiX7.cu:165-187)        */
    float    refine_distance; /* mm, PropagateRefineDistance, default 5000            */
    uint32_t vizmask;         /* ray visibility mask: an instance of solid s is seen
                                 when vizmask & (1 << min(s, 7)) (Params.cc:156,
                                 Properties.cc:45-55, SBT.cc:538-545); default 0xff  */
    int32_t  reserved;        /* 0                                                    */
} oxr_config;

/* Persistent-form launch parameters, chosen by the autotuner. */
typedef struct oxr_tuning {
    int32_t block_size;       /* lanes per workgroup (multiple of 64)                  */
    int32_t blocks_per_cu;    /* persistent grid = blocks_per_cu * num_CU              */
    int32_t lds_top_nodes;    /* top-level BVH nodes staged in LDS per workgroup       */
    int32_t refill_batch;     /* photons claimed per wave per queue visit (x64 lanes)  */
    int32_t stack_mode;       /* 1: traversal stack entries carry their box entry
                                 distance and popped nodes beyond the best hit are
                                 skipped (production kernels, when the block's LDS
                                 share at 4 waves/SIMD allows); 0: node indices only */
    int32_t bvh_mode;         /* 1: one BVH over every placed prim (world boxes), when the
                                 engine built it (at most 2^21 placed prims); 0: instance
                                 BVH over per-solid prim BVHs                            */
    int32_t csg_stack;        /* CSG evaluation stack depth: 1 keeps 3 partial intersects
                                 and 3 tranches in registers (production kernels, when every
                                 CSG tree has height 1 or 2 and no prim holds a list node);
                                 0 keeps the reference's 15 and 4 in private memory       */
    int32_t bvh_radial;       /* 1: BVH nodes also cull by radial bounds (a centre and the
                                 range of distances of each child's surfaces from it: exact
                                 for sphere prims, box-based otherwise), which removes the
                                 concentric shells whose boxes all hold the ray origin     */
    int32_t photon_order;     /* 1: production persistent launches claim photons sorted by
                                 (genstep, direction cell) after generation, so the lanes of
                                 a wave trace rays from one origin in neighbouring
                                 directions (results unchanged: every photon keeps its
                                 index and random stream); 0: index order                 */
    int32_t generate_inplace; /* 1: production launches generate each photon where the
                                 bounce loop claims it, with no generation pass before (so
                                 photon_order has nothing to sort and runs as 0); 0: a
                                 generation pass first, whose photons the loop resumes    */
} oxr_tuning;

/* ---- outputs ---------------------------------------------------------- */
typedef struct oxr_launch_stats {
    int64_t  num_photon;
    int64_t  num_hit;
    uint64_t bounce_steps;    /* sum over photons of loop iterations (trace+propagate) */
    double   kernel_ms;       /* simulate kernel, HIP events on the engine stream       */
    double   compact_ms;      /* hit compaction kernels                                 */
    double   total_ms;        /* launch call wall time                                  */
    uint64_t node_visits;     /* BVH node visits (when counting is enabled)             */
    uint64_t prim_tests;      /* CSG prim intersect calls (when counting is enabled)    */
    uint64_t csg_node_bytes;  /* 64 B per CSG node of each tested prim + 64 B per
                                 transformed leaf (when counting is enabled)            */
    uint64_t instance_visits; /* instance records read on entering a BLAS (counting)    */
    double   generate_ms;     /* the generation pass inside kernel_ms (production
                                 launches; 0 when debug outputs generate in place)     */
} oxr_launch_stats;

typedef struct oxr_engine_info {
    int64_t num_solid, num_prim, num_node, num_inst;
    int64_t tlas_nodes, blas_nodes;
    int32_t node_bytes;       /* BVH node size in bytes                                 */
    int32_t num_cu;
    int64_t device_bytes;     /* HBM held by geometry + tables                          */
    int64_t free_bytes, total_bytes;
    char    device_name[64];
    char    arch[32];
    int32_t csg_height;       /* greatest CSG tree height over the prims (0: leaves, lists)  */
    int32_t csg_in_registers; /* 1: production launches keep the CSG evaluation stacks in
                                 registers (tuning csg_stack 1 in force: every tree of height
                                 1 or 2, no list nodes)                                       */
    int32_t simple_leaves;    /* 1: every leaf is a sphere, zsphere, cylinder or box and no
                                 prim holds a list node, so production launches with single-
                                 precision roots run kernels compiled with those leaves only */
    int32_t csg_in_lds;       /* 1: those kernels, launched in the persistent form with the
                                 current tuning, copy the CSG node and transform arrays and
                                 the prim table into LDS and read them there (they take at
                                 most 16 KB and the copy leaves as many workgroups resident
                                 per CU)                                                    */
} oxr_engine_info;

typedef struct oxr_engine oxr_engine;
typedef struct oxr_server oxr_server;

/* ---- library ---------------------------------------------------------- */
const char* oxr_version(void);
const char* oxr_last_error(void);

/* ---- engine lifecycle --------------------------------------------------
 * Replaces CSGOptiX::Create(CSGFoundry*) (CSGOptiX/CSGOptiX.cc:367-399) together with
 * QSim::UploadComponents (qudarap/QSim.cc:118-250) and SBT::createGeom
 * (CSGOptiX/SBT.cc:247-259): copies geometry and tables to HBM and builds the BVHs. */
int  oxr_engine_create(const oxr_geometry* geo, const oxr_physics* phys,
                       const oxr_config* cfg, oxr_engine** out);
void oxr_engine_destroy(oxr_engine* e);
int  oxr_engine_set_config(oxr_engine* e, const oxr_config* cfg);
/* The persistent form's launch settings. Stored (and returned by get_tuning) as the
 * launches run them: stack_mode 1 falls back to 0 where the distance-carrying stack and
 * the staged top nodes do not fit a workgroup's LDS share, and lds_top_nodes is clipped to
 * what is staged. block_size is at most 256 (the kernels' launch bound). */
int  oxr_engine_set_tuning(oxr_engine* e, const oxr_tuning* t);
int  oxr_engine_get_tuning(oxr_engine* e, oxr_tuning* t);
/* The current stepping/event config (the SEventConfig getters, sysrap/SEventConfig.hh). */
int  oxr_engine_get_config(oxr_engine* e, oxr_config* cfg);
int  oxr_engine_query(oxr_engine* e, oxr_engine_info* info);
/* Fill *cfg with the reference defaults (SEventConfig.cc:45-106). */
void oxr_config_defaults(oxr_config* cfg);

/* ---- CSGFoundry folders -------------------------------------------------
 * Replaces CSGFoundry::Load (CSG/CSGFoundry.cc:2889-3020) with SSim::load
 * (sysrap/SSim.cc:486-520): reads <cfbase>/CSGFoundry (or cfbase itself when it holds
 * solid.npy) - solid, prim, node, plan, itra, inst .npy and SSim/stree/standard bnd,
 * optical, icdf .npy with the bnd_meta.txt wavelength domain - on the host. No device
 * calls. */
typedef struct oxr_foundry oxr_foundry;
int  oxr_foundry_load(const char* cfbase, oxr_foundry** out);
/* Views of the loaded arrays for oxr_engine_create, valid until oxr_foundry_free. */
int  oxr_foundry_arrays(const oxr_foundry* f, oxr_geometry* geo, oxr_physics* phys);
void oxr_foundry_free(oxr_foundry* f);

/* ---- per-event state ---------------------------------------------------
 * Replaces QEvt::setGenstepUpload_NP (qudarap/QEvt.cc:199-256,332-445) and
 * QEvt_count_genstep_photons_and_fill_seed_buffer (qudarap/QEvt.cu:181-237):
 * uploads gensteps, sums numphoton (gs.q0.u.w) and expands the per-photon
 * genstep index. photon_slot_offset keeps RNG subsequences absolute across
 * slices (CSGOptiX7.cu:415, QSim.cc:815). */
int oxr_event_set_gensteps(oxr_engine* e, const float* gs, int64_t num_gs,
                           int32_t event_index, uint64_t photon_slot_offset,
                           int64_t* num_photon);
/* INPUT_PHOTON source (qsim.h:2537): photons copied before launch, read by generate. */
int oxr_event_set_input_photons(oxr_engine* e, const float* photons, int64_t n);
/* Precooked per-photon uniform streams (n_photon, n_per_photon) f32, the format of
 * QSim::rng_sequence (qudarap/QSim.cc:1088) and sysrap/s_seq.h; NULL clears. Row i is
 * consumed by photon photon_slot_offset+i. */
int oxr_event_set_precooked(oxr_engine* e, const float* u, int64_t n_photon,
                            int32_t n_per_photon);

/* ---- launch ------------------------------------------------------------
 * Replaces SSimulator::simulate_launch (sysrap/SSimulator.h:36-55) as implemented by
 * CSGOptiX::launch (CSGOptiX/CSGOptiX.cc:1122-1198) followed by hit selection
 * SU::count_if/copy_if (sysrap/SU.cu:49-166). Synchronous. */
int oxr_simulate_launch(oxr_engine* e, int32_t form, oxr_launch_stats* stats);

/* Same launch as oxr_simulate_launch(OXR_FORM_PLAIN) with per-lane counting of BVH node
 * visits and CSG prim tests into stats (node_visits, prim_tests): the measured inputs of
 * the algorithmic-bytes model. No reference counterpart (instrumentation only). */
int oxr_simulate_launch_counting(oxr_engine* e, oxr_launch_stats* stats);

/* Gather: replaces QEvt::gatherPhoton/gatherHit/gatherSeq/gatherRecord
 * (qudarap/QEvt.cc:934-970,1113-1133). Copies min(count, capacity) items to host. */
int oxr_event_get(oxr_engine* e, int32_t what, void* dst, int64_t capacity,
                  int64_t* count);

/* Hit merging on the device: replaces QEvt::gatherHitMerged / gatherHitLiteMerged
 * (qudarap/QEvt.cc:1034-1086,1159-1195) and SPM::merge_partial_select (sysrap/SPM.cu:135-260).
 * The current launch's photons (what = OXR_GET_HIT, sphoton) or photonlite items
 * (what = OXR_GET_HITLITE, needs photonlite) whose flagmask shares any bit with the
 * config hitmask are keyed (identity<<48 | unsigned(time/time_window)), stably sorted by
 * key, and each run of equal keys is folded into one item: flagmasks OR-ed, 16-bit
 * hitcounts summed, sphoton keeping the earliest item, sphotonlite the first item with
 * the run's minimum time (reduce_op, sphoton.h:288-298, sphotonlite.h:79-90). Items come
 * out in key order. time_window 0 returns the selection in photon order, unmerged. */
int oxr_event_get_merged(oxr_engine* e, int32_t what, float time_window, void* dst,
                         int64_t capacity, int64_t* count);

/* Merge an already selected host array, e.g. the concatenated per-slice merges of a sliced
 * event: replaces QEvt::FinalMerge (qudarap/QEvt.cc:1236-1300, ALREADY_HITMASK_SELECTED). */
int oxr_merge_hits(oxr_engine* e, int32_t what, float time_window, const void* items,
                   int64_t n, void* dst, int64_t capacity, int64_t* count);

/* One request, one event: replaces CSGOptiXService::simulate(gs, eventID) -> hits
 * (CSGOptiX/CSGOptiXService.h:76-84, QSim::simulate qudarap/QSim.cc:711-740). Slices
 * the gensteps by config max_photon (QSim.cc:455-526) and concatenates hits. */
int oxr_simulate(oxr_engine* e, const float* gs, int64_t num_gs, int32_t event_index,
                 int32_t form, void* hits, int64_t hit_capacity, int64_t* num_hit,
                 oxr_launch_stats* stats);

/* One simulate launch that allocates no device memory. No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Host gensteps in (at most OXR_NOALLOC_MAX_GENSTEP,
 * at most OXR_NOALLOC_MAX_PHOTON photons in total, no INPUT_PHOTON gensteps); the caller's
 * device arrays out: d_photon (64 B per photon, in photon order) and d_hit (64 B per hit,
 * up to hit_capacity, the hits in photon order); *num_hit is the hit count
 * (OXR_ECAPACITY when it exceeds hit_capacity). Photon indices run from
 * photon_slot_offset, as in a slice of oxr_simulate. Production kernels only: the
 * config's seqhis, record, rec, prd and photonlite outputs are not written. Everything
 * else the launch needs lives in a library-owned device array of
 * OXR_NOALLOC_SCRATCH_BYTES on each GPU, held until the process ends and shared by
 * every engine on that GPU; calls within a process are serialised by a lock. Larger
 * events are refused with OXR_ECAPACITY. */
#define OXR_NOALLOC_MAX_PHOTON  (1 << 19)
#define OXR_NOALLOC_MAX_GENSTEP (1 << 14)
#define OXR_NOALLOC_SCRATCH_BYTES 12716288
int oxr_simulate_noalloc(oxr_engine* e, const float* gs, int64_t num_gs, int32_t event_index,
                         uint64_t photon_slot_offset, int32_t form, float* d_photon, float* d_hit,
                         int64_t hit_capacity, int64_t* num_hit, oxr_launch_stats* stats);

/* ---- geometry query ----------------------------------------------------
 * Closest hit through the two-level BVH for arbitrary rays: the trace() contract of
 * CSGOptiX7.cu:110-216 with __intersection__is/__closesthit__ch/__miss__ms
 * (CSGOptiX7.cu:655-940). rays (n,2,4) f32 = {pos.xyz, tmin}, {dir.xyz, tmax};
 * prd out (n,2,4) = quad2 {normal(world, unnormalised), t}, {lposcost, lposfphi,
 * iindex<<16|identity, globalPrimIdx<<16|boundary}; miss -> boundary 0xffff. */
int oxr_trace_rays(oxr_engine* e, const float* rays, int64_t n, float* prd);

/* ---- simtrace: 2-D geometry cross-sections -----------------------------
 * QSim::simtrace (qudarap/QSim.cc:857-890) launching the simtrace raygen
 * (CSGOptiX7.cu:536-577) over FRAME gensteps (gencode 17, q0 = {17, grid axes 0-3 =
 * XYZ/YZ/XZ/XY, genstep id, num rays}, q1 = local start position, q2..q5 = qat4
 * transform; SFrameGenstep::MakeCenterExtentGenstep, sysrap/SFrameGenstep.cc:604-700).
 * Each ray's direction comes from two uniforms of its photon-index stream
 * (qsim::generate_photon_simtrace_frame, qsim.h:2459-2510). Output (n,4,4) f32 as
 * sevent::add_simtrace (sysrap/sevent.h:670-696): {normal, t}, {pos + t dir, tmin},
 * {pos, globalPrimIdx<<16|boundary}, {dir, iindex<<16|identity}; a miss has t 0 and
 * 0xffffffff words. *num_out is the total ray count; OXR_ECAPACITY (nothing written)
 * when it exceeds capacity. */
int oxr_simtrace(oxr_engine* e, const float* gensteps, int64_t num_gensteps, float* simtrace,
                 int64_t capacity, int64_t* num_out);
/* Single-prim intersect_prim in the prim's own frame (CSG/csg_intersect_tree.h:684-719,
 * the CSGQuery::intersect contract CSG/CSGQuery.cc:231-276): rays (n,2,4) as above,
 * isect out (n,4) = normal.xyz (object frame), t; t=0 and valid=0 on miss. */
int oxr_intersect_prim(oxr_engine* e, int32_t prim_index, const float* rays, int64_t n,
                       float* isect, int32_t* valid);

/* ---- autotuner ----------------------------------------------------------
 * Times the persistent form on the current event over a grid of oxr_tuning values
 * (repeat times each), times the four fastest again in turn over three rounds (their
 * rows then report the mean of the rounds, as does any row that became the fastest
 * before it was timed so) and applies the fastest. results (max_results,OXR_AUTOTUNE_COLUMNS)
 * f64 = block_size, blocks_per_cu, lds_top_nodes, refill_batch, stack_mode, bvh_mode,
 * csg_stack, bvh_radial, photon_order, generate_inplace, kernel_ms,
 * one row per distinct effective tuning launched (see oxr_engine_set_tuning);
 * *num_results = rows written (at most max_results). On an error the tuning in force
 * before the call is restored. No reference counterpart. */
#define OXR_AUTOTUNE_COLUMNS 11
int oxr_autotune(oxr_engine* e, int32_t repeat, double* results, int32_t max_results,
                 int32_t* num_results);

/* ---- serving form --------------------------------------------------------
 * Continuous batching over one resident engine: requests (genstep batches of one
 * event each) are admitted while the GPU is busy as photon slots and free HBM allow
 * (64 B/photon state + hit buffer), their photons share persistent-kernel steps, and
 * each request completes independently. No reference counterpart beyond the one-at-a-
 * time request/response contract of CSGOptiXService (CSGOptiXService.h:76-84). */
typedef struct oxr_server_config {
    int64_t max_step_photons;   /* photons per batched device step                     */
    int64_t hbm_budget_bytes;   /* admission budget; 0 = 80% of free HBM at create      */
    float   hit_fraction;       /* hit buffer reserved per photon (fraction of 64 B)    */
    int32_t form;               /* OXR_FORM_PERSISTENT (default) or OXR_FORM_PLAIN      */
    int32_t max_active;         /* max concurrently admitted requests                   */
    int32_t headroom;           /* 0 (default): served persistent steps use the tuned    *
                                 * grid; 1: one workgroup per CU fewer than fit, leaving *
                                 * room for the other step slot's generate and           *
                                 * compaction kernels during a step                      */
} oxr_server_config;

typedef struct oxr_request_result {
    int64_t request_id;
    int32_t event_index;
    int32_t status;             /* 0 ok, else OXR_E* code                               */
    int64_t num_photon;
    int64_t num_hit;
    uint64_t bounce_steps;
    double  t_submit;           /* seconds on the server clock                          */
    double  t_admit;
    double  t_done;
} oxr_request_result;

int  oxr_server_create(oxr_engine* e, const oxr_server_config* cfg, oxr_server** out);
void oxr_server_destroy(oxr_server* s);
double oxr_server_clock(oxr_server* s);
/* Thread-safe; copies gensteps. Returns immediately. */
int  oxr_server_submit(oxr_server* s, int64_t request_id, const float* gs, int64_t num_gs,
                       int32_t event_index);
/* Wait for the next completed request (any id). timeout_ms < 0 waits forever. */
int  oxr_server_wait(oxr_server* s, oxr_request_result* res, int32_t timeout_ms);
/* Copy the hits of a completed request (sphoton, photon order) into dst (host memory)
 * and release them. Served hits cross PCIe once, inside the serving pipeline, into
 * pinned host memory the server owns; this call is a host copy and does no device work.
 * capacity below the count returns OXR_ECAPACITY with *count set, copies nothing and
 * keeps the request for a retry. */
int  oxr_server_take_hits(oxr_server* s, int64_t request_id, void* dst, int64_t capacity,
                          int64_t* count);
/* Aggregate, out (8) f64: steps run, photons, busy ms (wall time with at least one
 * step in flight), kernel ms (sum over steps of generate + propagate; steps overlap, so
 * this can exceed busy), bounce steps, hits, budget bytes in use, budget bytes. */
int  oxr_server_stats(oxr_server* s, double* out8);

#ifdef __cplusplus
}
#endif
#endif /* OXR_H */
