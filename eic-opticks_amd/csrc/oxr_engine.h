// oxr_engine.h -- host-side engine state shared by the C-ABI (oxr_engine.cpp) and the
// serving loop (oxr_server.cpp). The engine owns every device allocation it makes.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include <mutex>
#include "../../include/oxr.h"
#include "oxr_kernels.h"

namespace oxh {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define OXR_HIP(call)                                                        \
    do {                                                                     \
        hipError_t _e = (call);                                              \
        if (_e != hipSuccess) return ::oxh::hip_fail(_e, #call);             \
    } while (0)

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;     // capacity in elements
    hipError_t reserve(size_t want) {
        if (want <= n) return hipSuccess;
        if (p) { hipError_t e = hipFree(p); if (e != hipSuccess) return e; p = nullptr; n = 0; }
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), want * sizeof(T));
        if (e == hipSuccess) n = want;
        return e;
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
};

// One event slice's device buffers (the QEvt role).
struct EventBuffers {
    DBuf<float4> gs;           // 6 float4 per genstep
    DBuf<long long> gs_off;    // num_gs + 1 prefix offsets
    DBuf<int> seed;
    DBuf<float4> photon;       // 4 float4 per photon
    DBuf<ulonglong2> seq;      // 2 per photon
    DBuf<float4> record;
    DBuf<uint4> rec;           // srec per (photon, bounce) slot (config max_rec)
    DBuf<float4> prd;          // quad2 per (photon, bounce) slot (config max_prd)
    DBuf<uint8_t> hitflag;
    DBuf<float4> hit;
    DBuf<int> counts;
    DBuf<long long> offs;
    DBuf<unsigned long long> counters;
    DBuf<float4> input;
    DBuf<float> pre;
    DBuf<uint4> lite;          // sphotonlite per photon (config photonlite)
    DBuf<uint4> hitlite;       // selected sphotonlite, formed on demand
    DBuf<uint32_t> merged;     // merge output (sphoton or sphotonlite words)
    DBuf<ulonglong2> rstate;   // generate -> propagate hand-over, 16 B per photon
    DBuf<unsigned long long> okey;   // photon order (tuning photon_order): 2 keys per photon,
    DBuf<unsigned> oval;             // 2 indices per photon
    DBuf<char> otmp;                 // and the sort's work space
    void release_all();
};

struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    size_t total_mem = 0;   // device memory in bytes
    std::string name, arch;
    oxr_config cfg{};
    oxr_tuning tune{};
    // geometry
    DBuf<float4> node, plan, itra, bnd;
    DBuf<int2> prim;
    DBuf<oxd::BNode> tlas, blas;
    DBuf<oxd::InstD> inst;
    DBuf<int2> flat;             // (instance, local prim) per leaf of the flattened BVH
    DBuf<oxd::BNode> tlas_f;     // the flattened BVH (built when the placed prims are few enough)
    DBuf<float4> rad_tlas, rad_blas, rad_flat;   // radial bounds, two float4 per node of each tree
    bool has_flat = false;
    int tlas_root_f = 0;
    int64_t tlas_nodes_f = 0;
    int stack_need_f = 0;
    bool use_flat(const oxr_tuning& t) const { return has_flat && t.bvh_mode == 1; }
    bool use_flat() const { return use_flat(tune); }
    int64_t top_nodes(const oxr_tuning& t) const { return use_flat(t) ? tlas_nodes_f : tlas_nodes; }
    int64_t top_nodes() const { return top_nodes(tune); }
    DBuf<uint4> optical;
    DBuf<float> icdf;
    int tlas_root = 0;
    int num_inst = 0, num_solid = 0, num_prim = 0, num_node = 0, num_itra = 0;
    int64_t tlas_nodes = 0, blas_nodes = 0;
    int stack_need = 0;
    std::vector<int> solid_prim_offset, solid_num_prim, solid_blas_root;
    oxd::Tables tables{};
    int64_t geometry_bytes = 0;
    // current event (single-slice state, as QEvt)
    EventBuffers ev;
    std::vector<float> h_gs;
    int64_t num_gs = 0, num_photon = 0, num_hit = 0;
    int32_t event_index = 0;
    uint64_t slot_offset = 0;
    bool have_input = false, have_pre = false;
    int64_t input_n = 0;
    int32_t npre = 0;
    int64_t pre_rows = 0;
    int num_bnd = 0;
    bool has_list = false;   // some prim holds a list node: launches use the LIST kernels
    int csg_height = 0;      // greatest CSG tree height over the prims (0: leaves and lists only)
    bool simple_leaves = false;   // every leaf is a sphere, zsphere, cylinder or box
    // the register-resident CSG stacks can serve every tree (tuning csg_stack 1)
    bool registers_csg_ok() const { return !has_list && csg_height >= 1 && csg_height <= 2; }
    bool launched = false;
    bool hitlite_ready = false;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, eg = nullptr;
    // pinned staging for gathers (grown on demand up to 8 MB); see d2h
    void* h_stage = nullptr;
    size_t h_stage_bytes = 0;
    int d2h(void* dst, const void* src, size_t bytes);
    std::mutex mu;   // serialises C-ABI calls on one engine

    oxd::Geom geom() const;
    oxd::Params params(EventBuffers& b, int64_t n, int64_t slot_off, int32_t evt_idx, int block) const;
    int stack_cap(const oxr_tuning& t) const;
    int stack_cap() const { return stack_cap(tune); }
    int ntop_for(const oxr_tuning& t, int block) const;
    bool csg_copy_for(size_t smem, int block) const;
    bool tn_stack_for(const oxr_tuning& t, int block) const;
    // t as the launches of the persistent form run it: stack_mode 1 only where the
    // distance-carrying stack fits (tn_stack_for), lds_top_nodes as many as are staged
    oxr_tuning effective(const oxr_tuning& t) const;
};

// upload a gensteps slice into b, fill the seed array; returns photons in the slice
int upload_gensteps(Engine& e, EventBuffers& b, hipStream_t s, const float* gs, int64_t num_gs, int64_t* num_photon);
// run one simulate launch (+ compaction) on b for n photons; stats filled
int run_launch(Engine& e, EventBuffers& b, hipStream_t s, int64_t n, int64_t slot_off, int32_t evt_idx,
               int form, bool count, oxr_launch_stats* st, const float* pre, int npre, const float4* input);
int ensure_event_capacity(Engine& e, EventBuffers& b, int64_t n);

} // namespace oxh
