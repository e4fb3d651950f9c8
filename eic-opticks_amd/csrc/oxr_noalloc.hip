// oxr_noalloc.hip -- oxr_simulate_noalloc: one simulate launch that allocates no device
// memory. It has no counterpart in the reference; it is for callers that cannot allocate
// (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The caller passes the gensteps (host) and the photon and hit arrays (device). Every
// other buffer of the launch lives in `scratch`, a file-scope device array of fixed size
// that each GPU the library runs on holds until the process ends:
//   gensteps (96 B each) and their photon prefix offsets, the genstep index of each photon,
//   the generate -> propagate hand-over state (16 B per photon), the hit flags (1 B per
//   photon), the compaction counts and offsets (per 4096 photons) and the launch counters.
// Limits: OXR_NOALLOC_MAX_PHOTON photons and OXR_NOALLOC_MAX_GENSTEP gensteps per call.
// The array keeps the last call's contents; what must start at zero (the counters) is
// cleared by a fill kernel at the start of each call. Calls in one process take a lock, so
// they never overlap; calls from several processes on one GPU each have their own array.
#include "oxr_engine.h"
#include <algorithm>
#include <cstring>

namespace oxna {

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
constexpr int64_t MAX_PHOTON = OXR_NOALLOC_MAX_PHOTON;
constexpr int64_t MAX_GENSTEP = OXR_NOALLOC_MAX_GENSTEP;
constexpr int64_t CHUNK = 4096;                       // photons per compaction block
constexpr int64_t MAX_BLOCKS = (MAX_PHOTON + CHUNK - 1) / CHUNK;
// byte offsets of the regions of the scratch array
constexpr size_t OFF_GS = 0;
constexpr size_t OFF_GSOFF = align256(OFF_GS + size_t(MAX_GENSTEP) * 96);
constexpr size_t OFF_SEED = align256(OFF_GSOFF + size_t(MAX_GENSTEP + 1) * sizeof(long long));
constexpr size_t OFF_RSTATE = align256(OFF_SEED + size_t(MAX_PHOTON) * sizeof(int));
constexpr size_t OFF_FLAG = align256(OFF_RSTATE + size_t(MAX_PHOTON) * sizeof(ulonglong2));
constexpr size_t OFF_COUNTS = align256(OFF_FLAG + size_t(MAX_PHOTON));
constexpr size_t OFF_OFFS = align256(OFF_COUNTS + size_t(MAX_BLOCKS) * sizeof(int));
constexpr size_t OFF_CTR = align256(OFF_OFFS + size_t(MAX_BLOCKS + 1) * sizeof(long long));
constexpr size_t BYTES = align256(OFF_CTR + 8 * sizeof(unsigned long long));
static_assert(BYTES == OXR_NOALLOC_SCRATCH_BYTES, "oxr.h states the scratch size");

// the device address of `scratch` on each GPU (hipGetSymbolAddress), set once per device
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
unsigned char* g_scratch[MAX_DEVICES] = {};
long long g_offsets[MAX_GENSTEP + 1];                  // host staging of the prefix offsets

__global__ void k_zero_counters(unsigned long long* c) {
    if (threadIdx.x < 8) c[threadIdx.x] = 0ull;
}

} // namespace oxna

// Host code must not pass `scratch` itself to a kernel (in host code the name is a
// stand-in, not a device address): launches take the address hipGetSymbolAddress gave.
static __device__ __align__(256) unsigned char scratch[oxna::BYTES];

namespace oxl {

// Called by oxr_engine_create with the engine's device current: looks up this device's
// address of `scratch` once. Returns the HIP status of the lookup.
hipError_t noalloc_setup(int device) {
    if (device < 0 || device >= oxna::MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(oxna::g_mu);
    if (oxna::g_scratch[device]) return hipSuccess;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    oxna::g_scratch[device] = static_cast<unsigned char*>(p);
    return hipSuccess;
}

} // namespace oxl

using namespace oxh;

extern "C" int oxr_simulate_noalloc(oxr_engine* eh, const float* gs, int64_t num_gs, int32_t event_index,
                                    uint64_t photon_slot_offset, int32_t form, float* d_photon, float* d_hit,
                                    int64_t hit_capacity, int64_t* num_hit, oxr_launch_stats* stats) {
    using namespace oxna;
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !num_hit || num_gs < 0 || (num_gs > 0 && !gs) || hit_capacity < 0 || (hit_capacity > 0 && !d_hit))
        return fail(OXR_EINVAL, "oxr_simulate_noalloc: bad argument");
    if (form != OXR_FORM_PLAIN && form != OXR_FORM_PERSISTENT) return fail(OXR_EINVAL, "oxr_simulate_noalloc: unknown form");
    *num_hit = 0;
    if (num_gs > MAX_GENSTEP) {
        return fail(OXR_ECAPACITY, "oxr_simulate_noalloc: " + std::to_string(num_gs) + " gensteps exceed the scratch limit of " +
                                       std::to_string(MAX_GENSTEP));
    }
    std::lock_guard<std::mutex> lk(g_mu);
    std::lock_guard<std::mutex> elk(e->mu);
    unsigned char* S = (e->device >= 0 && e->device < MAX_DEVICES) ? g_scratch[e->device] : nullptr;
    if (!S) return fail(OXR_ESTATE, "oxr_simulate_noalloc: no scratch address for this device (setup failed)");
    // validate the gensteps and form the prefix offsets in the host staging array
    long long n = 0;
    g_offsets[0] = 0;
    for (int64_t g = 0; g < num_gs; g++) {
        const float* q = gs + 24 * g;
        uint32_t code, np, matline;
        std::memcpy(&code, q, 4);
        std::memcpy(&matline, q + 2, 4);
        std::memcpy(&np, q + 3, 4);
        if (code == 19u) return fail(OXR_EINVAL, "oxr_simulate_noalloc: INPUT_PHOTON gensteps are not accepted");
        if ((code == 15u || code == 18u || code == 1u) && int64_t(matline) >= 4 * int64_t(e->num_bnd))
            return fail(OXR_EINVAL, "oxr_simulate_noalloc: Cerenkov genstep matline outside the boundary table");
        n += (long long)np;
        if (n > MAX_PHOTON) {
            return fail(OXR_ECAPACITY, "oxr_simulate_noalloc: more than the scratch limit of " + std::to_string(MAX_PHOTON) +
                                           " photons");
        }
        g_offsets[g + 1] = n;
    }
    if (n > 0 && !d_photon) return fail(OXR_EINVAL, "oxr_simulate_noalloc: null photon array");
    int block = form == OXR_FORM_PERSISTENT ? e->tune.block_size : 256;
    if (block < 64 || block > 256 || block % 64) return fail(OXR_EINVAL, "tuning block_size must be a multiple of 64 in [64,256]");
    hipStream_t s = e->stream;
    float4* d_gs = reinterpret_cast<float4*>(S + OFF_GS);
    long long* d_off = reinterpret_cast<long long*>(S + OFF_GSOFF);
    int* d_seed = reinterpret_cast<int*>(S + OFF_SEED);
    ulonglong2* d_rstate = reinterpret_cast<ulonglong2*>(S + OFF_RSTATE);
    uint8_t* d_flag = S + OFF_FLAG;
    int* d_counts = reinterpret_cast<int*>(S + OFF_COUNTS);
    long long* d_offs = reinterpret_cast<long long*>(S + OFF_OFFS);
    unsigned long long* d_ctr = reinterpret_cast<unsigned long long*>(S + OFF_CTR);
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, d_ctr);
    OXR_HIP(hipGetLastError());
    if (num_gs > 0) {
        OXR_HIP(hipMemcpyAsync(d_gs, gs, size_t(num_gs) * 96, hipMemcpyHostToDevice, s));
        OXR_HIP(hipMemcpyAsync(d_off, g_offsets, size_t(num_gs + 1) * sizeof(long long), hipMemcpyHostToDevice, s));
        OXR_HIP(oxl::fill_seed(s, d_off, int(num_gs), d_seed, n));
    }
    long long nhit = 0;
    unsigned long long ctr[8] = {0};
    if (n > 0) {
        // production kernels only: the debug outputs of the engine's config are not written
        oxd::Params P = e->params(e->ev, n, int64_t(photon_slot_offset), event_index, block);
        P.gs = d_gs; P.seed = d_seed; P.input = nullptr; P.pre = nullptr; P.npre = 0;
        P.photon = reinterpret_cast<float4*>(d_photon);
        P.seq = nullptr; P.record = nullptr; P.rec = nullptr; P.prd = nullptr; P.lite = nullptr;
        P.bounce_out = nullptr; P.overflow_out = nullptr;
        P.rstate = d_rstate; P.hitflag = d_flag; P.counters = d_ctr;
        P.segs = nullptr; P.nseg = 0;
        if (oxl::simulate_smem(P, block) > 64 * 1024) return fail(OXR_EINVAL, "LDS budget exceeded: lower block_size or lds_top_nodes");
        OXR_HIP(oxl::generate(s, P));
        if (form == OXR_FORM_PERSISTENT) {
            int grid = std::max(1, e->tune.blocks_per_cu) * std::max(1, e->num_cu);
            OXR_HIP(oxl::simulate_persist(s, P, block, grid, e->cfg.fp64_roots != 0, e->has_list));
        } else {
            OXR_HIP(oxl::simulate_plain(s, P, block, e->cfg.fp64_roots != 0, false, e->has_list));
        }
        OXR_HIP(oxl::compact(s, d_flag, P.photon, n, d_counts, d_offs, reinterpret_cast<float4*>(d_hit), hit_capacity));
        OXR_HIP(hipMemcpyAsync(&nhit, d_offs + oxl::compact_blocks(n), sizeof(long long), hipMemcpyDeviceToHost, s));
        OXR_HIP(hipMemcpyAsync(ctr, d_ctr, sizeof(ctr), hipMemcpyDeviceToHost, s));
    }
    OXR_HIP(hipStreamSynchronize(s));
    if (ctr[4] != 0) return fail(OXR_EINVAL, "traversal stack overflow in BVH (geometry deeper than the stack)");
    *num_hit = nhit;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->num_photon = n;
        stats->num_hit = nhit;
        stats->bounce_steps = ctr[0];
    }
    if (nhit > hit_capacity) return fail(OXR_ECAPACITY, "hit array smaller than the hit count");
    return OXR_OK;
}
