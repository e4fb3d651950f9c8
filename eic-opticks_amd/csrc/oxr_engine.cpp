// oxr_engine.cpp -- C-ABI implementation: geometry/table upload with BVH builds,
// per-event genstep upload and seeding, launches, hit compaction, gathers, queries and
// the autotuner. The serving loop lives in oxr_server.cpp.
#include "oxr_engine.h"
#include "oxr_bvh.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <sstream>

namespace oxh {

static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
int fail(int code, const std::string& m) { g_err = m; return code; }
int hip_fail(hipError_t e, const char* what) {
    std::ostringstream ss;
    ss << "HIP error " << int(e) << " (" << hipGetErrorString(e) << ") in " << what;
    g_err = ss.str();
    return OXR_EHIP;
}

void EventBuffers::release_all() {
    gs.release(); gs_off.release(); seed.release(); photon.release(); seq.release(); record.release();
    rec.release(); prd.release();
    hitflag.release(); hit.release(); counts.release(); offs.release(); counters.release();
    input.release(); pre.release(); lite.release(); hitlite.release(); merged.release(); rstate.release();
    okey.release(); oval.release(); otmp.release();
}

static inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

oxd::Geom Engine::geom() const {
    oxd::Geom g;
    g.node = node.p; g.plan = plan.p; g.itra = itra.p; g.prim = prim.p;
    const bool f = use_flat();
    g.tlas = f ? tlas_f.p : tlas.p; g.blas = blas.p; g.inst = inst.p;
    g.flat = f ? flat.p : nullptr;
    g.tlas_root = f ? tlas_root_f : tlas_root; g.num_inst = num_inst;
    g.vizmask = cfg.vizmask;
    g.rad_top = f ? rad_flat.p : rad_tlas.p; g.rad_blas = rad_blas.p;
    return g;
}

int Engine::stack_cap(const oxr_tuning& t) const { return std::max(4, (use_flat(t) ? stack_need_f : stack_need) + 2); }

int Engine::ntop_for(const oxr_tuning& t, int block) const {
    int want = std::max(0, t.lds_top_nodes);
    want = int(std::min<int64_t>(want, top_nodes(t)));
    size_t lim = 64 * 1024;
    size_t stack = size_t(block) * stack_cap(t) * sizeof(int);
    if (stack >= lim) return 0;
    size_t room = (lim - stack) / sizeof(oxd::BNode);
    return int(std::min<size_t>(size_t(want), room));
}

// The distance-carrying stack (8 B per entry) is used when it fits, with the wanted top
// nodes, in a block's share of the CU's 160 KB of LDS at 4 waves per SIMD (16 waves, 1024
// lanes per CU): more LDS per block would cost occupancy, which the latency-bound
// traversal needs more than the culled pops.
bool Engine::tn_stack_for(const oxr_tuning& t, int block) const {
    int want = int(std::min<int64_t>(std::max(0, t.lds_top_nodes), top_nodes(t)));
    size_t share = size_t(160 * 1024) * size_t(block) / 1024;
    size_t need = size_t(want) * sizeof(oxd::BNode) + size_t(block) * stack_cap(t) * 8;
    return need <= share;
}

// The simple-leaf kernels read the node and itra arrays and the prim table from a copy in
// LDS when they take at most 16 KB and the copy leaves as many blocks of `block` lanes resident per CU
// as the launch has without it (smem bytes): 160 KB of LDS per CU, and at most 16 waves
// per CU at the kernels' 4 waves per SIMD. A copy that cost a resident block would cost
// more in latency hiding than its LDS reads save.
bool Engine::csg_copy_for(size_t smem, int block) const {
    const size_t bytes = size_t(num_node + num_itra) * 64 + size_t(num_prim) * 8;
    if (bytes > 16 * 1024 || smem + bytes > 64 * 1024 || block < 64) return false;
    const size_t waves = size_t(16 / (block / 64));
    auto resident = [&](size_t sm) { return std::min(waves, size_t(160 * 1024) / std::max<size_t>(sm, 1)); };
    return resident(smem + bytes) >= resident(smem);
}

oxr_tuning Engine::effective(const oxr_tuning& t) const {
    oxr_tuning r = t;
    r.bvh_mode = use_flat(t) ? 1 : 0;
    const bool tn = t.stack_mode == 1 && tn_stack_for(t, t.block_size);
    r.stack_mode = tn ? 1 : 0;
    r.lds_top_nodes = tn ? int(std::min<int64_t>(std::max(0, t.lds_top_nodes), top_nodes(t))) : ntop_for(t, t.block_size);
    r.csg_stack = t.csg_stack == 1 && registers_csg_ok() ? 1 : 0;
    r.photon_order = t.photon_order == 1 && t.generate_inplace != 1 ? 1 : 0;   // nothing to sort in place
    return r;
}

oxd::Params Engine::params(EventBuffers& b, int64_t n, int64_t slot_off, int32_t evt_idx, int block) const {
    const bool tn = tune.stack_mode == 1 && tn_stack_for(tune, block);
    const int ntop = tn ? int(std::min<int64_t>(std::max(0, tune.lds_top_nodes), top_nodes())) : ntop_for(tune, block);
    oxd::Params P{};
    P.g = geom();
    P.t = tables;
    P.gs = b.gs.p; P.seed = b.seed.p;
    P.input = nullptr; P.pre = nullptr; P.npre = 0;
    P.photon = b.photon.p;
    P.seq = cfg.seq_enabled ? b.seq.p : nullptr;
    P.record = cfg.max_record > 0 ? b.record.p : nullptr;
    P.max_record = cfg.max_record;
    P.rec = cfg.max_rec > 0 ? b.rec.p : nullptr;
    P.max_rec = cfg.max_rec;
    P.prd = cfg.max_prd > 0 ? b.prd.p : nullptr;
    P.max_prd = cfg.max_prd;
    P.extent_domain = cfg.extent_domain; P.time_domain = cfg.time_domain;
    P.lite = cfg.photonlite ? b.lite.p : nullptr;
    P.rstate = b.rstate.p;
    P.hitflag = b.hitflag.p;
    P.bounce_out = nullptr; P.overflow_out = nullptr;
    P.counters = b.counters.p;
    P.num_photon = n;
    P.slot_offset = uint64_t(slot_off);
    P.event_index = evt_idx;
    P.max_bounce = cfg.max_bounce;
    P.tmin = cfg.tmin; P.tmin0 = cfg.tmin0; P.eps0_mask = cfg.eps0_mask;
    P.tmax = cfg.tmax; P.max_time = cfg.max_time; P.hitmask = cfg.hitmask;
    P.rng_seed = cfg.rng_seed; P.rng_offset = cfg.rng_offset; P.skip = cfg.skipahead_event;
    P.aligned = cfg.rng_schedule == OXR_RNG_ALIGNED ? 1 : 0;
    P.ntop = ntop;
    P.stack_cap = stack_cap();
    P.tn_stack = tn ? 1 : 0;
    P.shallow_csg = tune.csg_stack == 1 && registers_csg_ok() ? 1 : 0;
    P.simple_leaves = simple_leaves && !has_list ? 1 : 0;
    P.csg_node4 = P.csg_itra4 = P.csg_prim = 0;
    if (P.simple_leaves && csg_copy_for(oxl::simulate_smem(P, block), block)) {
        P.csg_node4 = num_node * 4;
        P.csg_itra4 = num_itra * 4;
        P.csg_prim = num_prim;
    }
    P.gen_inplace = tune.generate_inplace == 1 ? 1 : 0;
    P.radial = tune.bvh_radial == 1 ? 1 : 0;
    P.refine = cfg.propagate_refine ? 1 : 0;
    P.refine_distance = cfg.refine_distance;
    P.refill_threshold = std::max(1, tune.refill_batch);
    P.segs = nullptr;
    P.nseg = 0;
    return P;
}

int ensure_event_capacity(Engine& e, EventBuffers& b, int64_t n) {
    size_t N = size_t(std::max<int64_t>(n, 1));
    OXR_HIP(b.seed.reserve(N));
    OXR_HIP(b.photon.reserve(4 * N));
    if (e.cfg.seq_enabled) OXR_HIP(b.seq.reserve(2 * N));
    if (e.cfg.max_record > 0) OXR_HIP(b.record.reserve(4 * N * size_t(e.cfg.max_record)));
    if (e.cfg.max_rec > 0) OXR_HIP(b.rec.reserve(N * size_t(e.cfg.max_rec)));
    if (e.cfg.max_prd > 0) OXR_HIP(b.prd.reserve(2 * N * size_t(e.cfg.max_prd)));
    if (e.cfg.photonlite) OXR_HIP(b.lite.reserve(N));
    OXR_HIP(b.hitflag.reserve(N));
    OXR_HIP(b.rstate.reserve(N));
    OXR_HIP(b.hit.reserve(4 * N));
    int nb = oxl::compact_blocks(int64_t(N));
    OXR_HIP(b.counts.reserve(size_t(nb)));
    OXR_HIP(b.offs.reserve(size_t(nb) + 1));
    OXR_HIP(b.counters.reserve(8));
    return OXR_OK;
}

// validate gensteps against the engine's tables and form the photon prefix offsets
int genstep_offsets(Engine& e, const float* gs, int64_t num_gs, std::vector<long long>& off, int64_t* num_photon) {
    if (num_gs < 0 || (num_gs > 0 && !gs)) return fail(OXR_EINVAL, "gensteps: null pointer or negative count");
    off.assign(size_t(num_gs) + 1, 0);
    for (int64_t g = 0; g < num_gs; g++) {
        const float* q = gs + 24 * g;
        uint32_t code = fbits(q[0]);
        uint32_t np = fbits(q[3]);
        if (code == 15 || code == 18 || code == 1) {
            uint32_t matline = fbits(q[2]);
            if (int64_t(matline) >= 4 * int64_t(e.num_bnd))
                return fail(OXR_EINVAL, "Cerenkov genstep matline outside the boundary table");
        }
        off[size_t(g) + 1] = off[size_t(g)] + (long long)np;
    }
    *num_photon = off[size_t(num_gs)];
    return OXR_OK;
}

int upload_gensteps(Engine& e, EventBuffers& b, hipStream_t s, const float* gs, int64_t num_gs, int64_t* num_photon) {
    std::vector<long long> off;
    int64_t n = 0;
    int rc0 = genstep_offsets(e, gs, num_gs, off, &n);
    if (rc0) return rc0;
    if (n > e.cfg.max_photon) {
        std::ostringstream ss;
        ss << "slice of " << n << " photons exceeds max_photon " << e.cfg.max_photon;
        return fail(OXR_EINVAL, ss.str());
    }
    int rc = ensure_event_capacity(e, b, n);
    if (rc) return rc;
    OXR_HIP(b.gs.reserve(size_t(std::max<int64_t>(num_gs, 1)) * 6));
    OXR_HIP(b.gs_off.reserve(size_t(num_gs) + 1));
    if (num_gs > 0) {
        OXR_HIP(hipMemcpyAsync(b.gs.p, gs, size_t(num_gs) * 96, hipMemcpyHostToDevice, s));
        OXR_HIP(hipMemcpyAsync(b.gs_off.p, off.data(), (size_t(num_gs) + 1) * sizeof(long long), hipMemcpyHostToDevice, s));
        OXR_HIP(oxl::fill_seed(s, b.gs_off.p, int(num_gs), b.seed.p, n));
    }
    OXR_HIP(hipStreamSynchronize(s));
    *num_photon = n;
    return OXR_OK;
}

int run_launch(Engine& e, EventBuffers& b, hipStream_t s, int64_t n, int64_t slot_off, int32_t evt_idx,
               int form, bool count, oxr_launch_stats* st, const float* pre, int npre, const float4* input) {
    auto t0 = std::chrono::steady_clock::now();
    int block = form == OXR_FORM_PERSISTENT ? e.tune.block_size : 256;
    if (block < 64 || block > 256 || block % 64) return fail(OXR_EINVAL, "tuning block_size must be a multiple of 64 in [64,256]");
    oxd::Params P = e.params(b, n, slot_off, evt_idx, block);
    P.pre = pre; P.npre = npre; P.input = input;
    size_t smem = oxl::simulate_smem(P, block);
    if (smem > 64 * 1024) return fail(OXR_EINVAL, "LDS budget exceeded: lower block_size or lds_top_nodes");
    OXR_HIP(hipMemsetAsync(b.counters.p, 0, 8 * sizeof(unsigned long long), s));
    // record slots past a photon's last point stay zero, as the reference's zeroed record array
    if (P.record && n > 0) OXR_HIP(hipMemsetAsync(P.record, 0, size_t(n) * size_t(e.cfg.max_record) * 64, s));
    if (P.rec && n > 0) OXR_HIP(hipMemsetAsync(P.rec, 0, size_t(n) * size_t(e.cfg.max_rec) * 16, s));
    if (P.prd && n > 0) OXR_HIP(hipMemsetAsync(P.prd, 0, size_t(n) * size_t(e.cfg.max_prd) * 32, s));
    OXR_HIP(hipEventRecord(e.e0, s));
    bool gen = oxl::needs_generate(P, count);
    if (gen) OXR_HIP(oxl::generate(s, P));
    OXR_HIP(hipEventRecord(e.eg, s));
    if (form == OXR_FORM_PERSISTENT) {
        int grid = std::max(1, e.tune.blocks_per_cu) * std::max(1, e.num_cu);
        if (e.tune.photon_order == 1 && gen && !count && P.seed && n > 1 && oxl::production_kernels(P)) {
            // claim photons sorted by (genstep, direction cell): keys of 16 direction bits
            // under as many genstep-index bits as the event needs
            int bits = 16;
            while (bits < 64 && (int64_t(1) << (bits - 16)) < e.num_gs) bits++;
            OXR_HIP(b.okey.reserve(2 * size_t(n)));
            OXR_HIP(b.oval.reserve(2 * size_t(n)));
            size_t tb = 0;
            OXR_HIP(oxl::order_photons(s, P, bits, b.okey.p, b.oval.p, nullptr, &tb, nullptr));
            OXR_HIP(b.otmp.reserve(std::max<size_t>(tb, 1)));
            OXR_HIP(oxl::order_photons(s, P, bits, b.okey.p, b.oval.p, b.otmp.p, &tb, &P.perm));
        }
        OXR_HIP(oxl::simulate_persist(s, P, block, grid, e.cfg.fp64_roots != 0, e.has_list));
    } else {
        OXR_HIP(oxl::simulate_plain(s, P, block, e.cfg.fp64_roots != 0, count, e.has_list));
    }
    OXR_HIP(hipEventRecord(e.e1, s));
    OXR_HIP(oxl::compact(s, b.hitflag.p, b.photon.p, n, b.counts.p, b.offs.p, b.hit.p, n));
    OXR_HIP(hipEventRecord(e.e2, s));
    unsigned long long ctr[8] = {0};
    long long nhit = 0;
    OXR_HIP(hipMemcpyAsync(ctr, b.counters.p, sizeof(ctr), hipMemcpyDeviceToHost, s));
    if (n > 0) OXR_HIP(hipMemcpyAsync(&nhit, b.offs.p + oxl::compact_blocks(n), sizeof(long long), hipMemcpyDeviceToHost, s));
    OXR_HIP(hipStreamSynchronize(s));
    if (ctr[4] != 0) return fail(OXR_EINVAL, "traversal stack overflow in BVH (geometry deeper than the stack)");
    float k_ms = 0.f, c_ms = 0.f;
    OXR_HIP(hipEventElapsedTime(&k_ms, e.e0, e.e1));
    float g_ms = 0.f;
    OXR_HIP(hipEventElapsedTime(&g_ms, e.e0, e.eg));
    OXR_HIP(hipEventElapsedTime(&c_ms, e.e1, e.e2));
    auto t1 = std::chrono::steady_clock::now();
    if (st) {
        st->num_photon = n;
        st->num_hit = nhit;
        st->bounce_steps = ctr[0];
        st->kernel_ms = k_ms;
        st->generate_ms = g_ms;
        st->compact_ms = c_ms;
        st->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st->node_visits = ctr[2];
        st->prim_tests = ctr[3];
        st->csg_node_bytes = ctr[5];
        st->instance_visits = ctr[6];
    }
    e.num_hit = nhit;
    e.hitlite_ready = false;
    return OXR_OK;
}

// inverse of the instance transform in double (row convention, identity column ignored)
static void invert_instance(const float* q, double W[4][3]) {
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
    for (int i = 0; i < 3; i++) for (int k = 0; k < 3; k++) W[i][k] = inv[i][k];
    for (int k = 0; k < 3; k++) W[3][k] = -(t[0] * inv[0][k] + t[1] * inv[1][k] + t[2] * inv[2][k]);
}

static bool supported_leaf(uint32_t tc) {
    return tc == 101 || tc == 103 || tc == 105 || tc == 108 || tc == 110 || tc == 112 || tc == 117 || tc == 121 ||
           tc == 125;
}

static bool is_list(uint32_t tc) { return tc == 11 || tc == 12 || tc == 13; }
static bool is_op(uint32_t tc) { return tc == 1 || tc == 2 || tc == 3; }

// Checks every prim against what the kernels assume, so no launch can index outside the
// arrays: node ranges, tree shape (root subNum = 2^k-1 tree nodes, operators above the
// leaf level), list nodes (CSG_CONTIGUOUS/DISCONTIGUOUS/OVERLAP: subs inside the prim and
// all leaves, at most 8 for contiguous), leaf typecodes, transform and plane ranges,
// boundary indices. Sets e.has_list when any prim holds a list node, e.csg_height to
// the greatest tree height and e.simple_leaves when every leaf is a sphere, zsphere,
// cylinder or box.
static int validate_and_prims(Engine& e, const oxr_geometry* g, const oxr_physics* ph, std::vector<int2>& primtab) {
    const uint32_t* nu = reinterpret_cast<const uint32_t*>(g->node);
    auto node_u = [&](int64_t i, int k) { return nu[16 * i + k]; };
    primtab.resize(size_t(g->num_prim));
    e.has_list = false;
    e.csg_height = 0;
    e.simple_leaves = true;
    for (int64_t p = 0; p < g->num_prim; p++) {
        const int32_t* pi = reinterpret_cast<const int32_t*>(g->prim + 16 * p);
        const uint32_t* pu = reinterpret_cast<const uint32_t*>(g->prim + 16 * p);
        int32_t numNode = pi[0], nodeOffset = pi[1];
        std::string tag = "prim " + std::to_string(p) + ": ";
        if (numNode < 1 || nodeOffset < 0 || int64_t(nodeOffset) + numNode > g->num_node)
            return fail(OXR_EINVAL, tag + "node range outside node array");
        auto check_leaf = [&](int64_t ni) -> int {
            uint32_t tc = node_u(ni, 14);
            if (!supported_leaf(tc)) return fail(OXR_EINVAL, tag + "unsupported leaf typecode " + std::to_string(tc));
            if (tc != 101 && tc != 103 && tc != 105 && tc != 110) e.simple_leaves = false;
            uint32_t gt = node_u(ni, 15) & 0x7fffffffu;
            if (int64_t(gt) > g->num_itra) return fail(OXR_EINVAL, tag + "node transform index outside itra");
            if (tc == 112) {
                uint32_t pidx = node_u(ni, 0), pnum = node_u(ni, 1);
                if (int64_t(pidx) + int64_t(pnum) > g->num_plan) return fail(OXR_EINVAL, tag + "convexpolyhedron planes outside plan array");
            }
            return OXR_OK;
        };
        auto check_list = [&](int64_t ni) -> int {
            uint32_t tc = node_u(ni, 14), num_sub = node_u(ni, 0), off_sub = node_u(ni, 1);
            if (num_sub < 1 || int64_t(off_sub) + int64_t(num_sub) > numNode)
                return fail(OXR_EINVAL, tag + "list node subs outside the prim's nodes");
            if (tc == 11 && num_sub > 8) return fail(OXR_EINVAL, tag + "CSG_CONTIGUOUS with more than 8 subs");
            for (uint32_t k = 0; k < num_sub; k++) {
                int rc = check_leaf(nodeOffset + int64_t(off_sub) + k);
                if (rc) return rc;
            }
            e.has_list = true;
            return OXR_OK;
        };
        uint32_t rtc = node_u(nodeOffset, 14);
        int rc = OXR_OK;
        if (rtc >= 101) {
            rc = check_leaf(nodeOffset);
        } else if (is_list(rtc)) {
            rc = check_list(nodeOffset);
        } else if (is_op(rtc)) {
            uint32_t T = node_u(nodeOffset, 0);
            if (T < 3 || T > 255 || ((T + 1) & T) != 0 || int64_t(T) > numNode)
                return fail(OXR_EINVAL, tag + "tree root subNum must be 2^k-1 <= 255 and within numNode");
            int height = 0;
            while ((1 << (height + 1)) - 1 < int(T)) height++;
            e.csg_height = std::max(e.csg_height, height);
            for (int i = 1; i <= int(T) && rc == OXR_OK; i++) {
                int64_t ni = nodeOffset + i - 1;
                uint32_t tc = node_u(ni, 14);
                int depth = 31 - __builtin_clz(unsigned(i));
                if (is_op(tc)) {
                    if (depth == height) rc = fail(OXR_EINVAL, tag + "operator at leaf level");
                } else if (is_list(tc)) {
                    rc = check_list(ni);
                } else if (tc != 0) {
                    rc = check_leaf(ni);
                }
            }
        } else {
            return fail(OXR_EINVAL, tag + "unsupported root typecode " + std::to_string(rtc));
        }
        if (rc) return rc;
        uint32_t bnd = node_u(nodeOffset, 6);
        if (int64_t(bnd) >= int64_t(ph->num_bnd)) return fail(OXR_EINVAL, tag + "boundary outside bnd table");
        uint32_t gp = pu[15];
        primtab[size_t(p)] = make_int2(nodeOffset, int(((gp & 0xffffu) << 16) | (bnd & 0xffffu)));
    }
    return OXR_OK;
}

} // namespace oxh

using namespace oxh;

extern "C" {

const char* oxr_version(void) { return "oxr 0.1 gfx950"; }
const char* oxr_last_error(void) { return g_err.c_str(); }

// The simulate kernels are compiled for at most 256 lanes per workgroup
// (__launch_bounds__(256, N)): a larger block would fail at launch.
static bool tuning_ok(const oxr_tuning& t) {
    return t.block_size >= 64 && t.block_size <= 256 && t.block_size % 64 == 0 && t.blocks_per_cu >= 1 &&
           t.blocks_per_cu <= 64 && t.lds_top_nodes >= 0 && t.refill_batch >= 1 && t.refill_batch <= 64 &&
           t.stack_mode >= 0 && t.stack_mode <= 1 && t.bvh_mode >= 0 && t.bvh_mode <= 1 && t.csg_stack >= 0 &&
           t.csg_stack <= 1 && t.bvh_radial >= 0 && t.bvh_radial <= 1 && t.photon_order >= 0 && t.photon_order <= 1 &&
           t.generate_inplace >= 0 && t.generate_inplace <= 1;
}

// SEventConfig::HeuristicMaxSlot (sysrap/SEventConfig.cc:1897-1919): with max_photon 0 a
// launch slice takes as many photons as fit in 87% of the device memory at this config's
// per-photon bytes (photon, genstep index, resume state, hit flag, a hit, and the debug
// arrays the config enables), rounded down to a million.
static int64_t heuristic_max_photon(const oxr_config& c, size_t total_bytes) {
    double per = 64.0 + 4.0 + 16.0 + 1.0 + 64.0;
    if (c.seq_enabled) per += 32.0;
    if (c.photonlite) per += 32.0;
    per += 64.0 * c.max_record + 16.0 * c.max_rec + 32.0 * c.max_prd;
    const int64_t n = int64_t(0.87 * double(total_bytes) / per / 1e6) * 1000000;
    return std::max<int64_t>(n, 1000000);
}

static bool config_ok(const oxr_config& c) {
    if (c.max_record < 0 || c.max_photon < 0 || c.max_rec < 0 || c.max_prd < 0) return false;
    // the rec domains divide each compressed coordinate
    if (c.max_rec > 0 && !(c.extent_domain > 0.f && c.time_domain > 0.f)) return false;
    if (c.propagate_refine > 1u || !std::isfinite(c.refine_distance)) return false;
    return true;
}

void oxr_config_defaults(oxr_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->max_bounce = 31;
    c->tmin = 0.05f; c->tmin0 = 0.05f;
    c->eps0_mask = (1u << 2) | (1u << 0) | (1u << 1) | (1u << 5) | (1u << 4);
    c->tmax = 1e6f; c->max_time = 1e27f;
    c->hitmask = 1u << 6;
    c->rng_seed = 0; c->rng_offset = 0; c->skipahead_event = 100000;
    c->rng_schedule = OXR_RNG_PRODUCTION;
    c->fp64_roots = 0; c->seq_enabled = 0; c->max_record = 0;
    c->max_photon = 0;   // MaxSlot 0: sized from the device memory (SEventConfig.cc:90,1838-1852)
    c->photonlite = 0;
    c->max_rec = 0; c->max_prd = 0;
    c->extent_domain = 1000.f; c->time_domain = 10.f;   // SEventConfig.cc:72-73
    c->propagate_refine = 0u; c->refine_distance = 5000.f;   // SEventConfig.cc:105-106
    c->vizmask = 0xffu;                                    // Params.cc:156
    c->reserved = 0;
}

int oxr_engine_create(const oxr_geometry* g, const oxr_physics* ph, const oxr_config* cfg, oxr_engine** out) {
    if (!g || !ph || !cfg || !out) return fail(OXR_EINVAL, "null argument");
    *out = nullptr;
    if (!g->solid || !g->prim || !g->node || !g->inst || g->num_solid < 1 || g->num_prim < 1 || g->num_node < 1 || g->num_inst < 1)
        return fail(OXR_EINVAL, "geometry needs solid, prim, node and inst arrays");
    if (g->num_itra > 0 && !g->itra) return fail(OXR_EINVAL, "itra pointer null");
    if (g->num_plan > 0 && !g->plan) return fail(OXR_EINVAL, "plan pointer null");
    if (!ph->bnd || !ph->optical || ph->num_bnd < 1 || ph->num_wl < 2 || !(ph->wl_step > 0.f))
        return fail(OXR_EINVAL, "physics needs bnd and optical tables");
    if (ph->icdf && (ph->icdf_rows < 1 || ph->icdf_width < 2)) return fail(OXR_EINVAL, "icdf shape");
    if (ph->icdf && ph->hd_factor != 0 && ph->icdf_rows < 3) return fail(OXR_EINVAL, "hd_factor needs 3 icdf rows");
    if (!config_ok(*cfg)) return fail(OXR_EINVAL, "config");
    int dev = 0;
    int ndev = 0;
    OXR_HIP(hipGetDeviceCount(&ndev));
    if (ndev < 1) return fail(OXR_EHIP, "no HIP device visible");
    OXR_HIP(hipGetDevice(&dev));
    Engine* e = new Engine();
    e->device = dev;
    e->cfg = *cfg;
    hipDeviceProp_t prop;
    hipError_t he = hipGetDeviceProperties(&prop, dev);
    if (he != hipSuccess) { delete e; return hip_fail(he, "hipGetDeviceProperties"); }
    e->num_cu = prop.multiProcessorCount;
    e->total_mem = prop.totalGlobalMem;
    if (e->cfg.max_photon == 0) e->cfg.max_photon = heuristic_max_photon(e->cfg, e->total_mem);
    e->name = prop.name;
    e->arch = prop.gcnArchName;
    // this device's address of the no-alloc entry point's scratch array, looked up once per
    // device; a failure here is reported by oxr_simulate_noalloc, not by the engine
    if ((he = oxl::noalloc_setup(dev)) != hipSuccess) set_error(std::string("no-alloc scratch: ") + hipGetErrorString(he));
    e->tune.block_size = 256;
    e->tune.blocks_per_cu = 3;
    e->tune.lds_top_nodes = 64;
    e->tune.refill_batch = 16;
    e->tune.stack_mode = 1;
    e->tune.bvh_mode = 0;
    e->tune.csg_stack = 0;
    e->tune.bvh_radial = 0;
    e->tune.photon_order = 0;
    e->tune.generate_inplace = 0;
    auto bail = [&](int rc) { oxr_engine_destroy(reinterpret_cast<oxr_engine*>(e)); return rc; };
    if ((he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess) return bail(hip_fail(he, "hipStreamCreate"));
    if ((he = hipEventCreate(&e->e0)) != hipSuccess) return bail(hip_fail(he, "hipEventCreate"));
    if ((he = hipEventCreate(&e->e1)) != hipSuccess) return bail(hip_fail(he, "hipEventCreate"));
    if ((he = hipEventCreate(&e->e2)) != hipSuccess) return bail(hip_fail(he, "hipEventCreate"));
    if ((he = hipEventCreate(&e->eg)) != hipSuccess) return bail(hip_fail(he, "hipEventCreate"));

    std::vector<int2> primtab;
    int rc = validate_and_prims(*e, g, ph, primtab);
    if (rc) return bail(rc);
    e->num_prim = int(g->num_prim); e->num_node = int(g->num_node); e->num_solid = int(g->num_solid);
    e->num_itra = int(g->num_itra);
    e->num_inst = int(g->num_inst);
    e->num_bnd = ph->num_bnd;

    // ---- per-solid BLAS over prim AABBs
    std::vector<oxh::NodeOut> blas_all;
    std::vector<oxh::RadOut> rad_blas;
    int max_blas_depth = 0;
    e->solid_prim_offset.resize(size_t(g->num_solid));
    e->solid_num_prim.resize(size_t(g->num_solid));
    e->solid_blas_root.resize(size_t(g->num_solid));
    std::vector<oxh::Box> solid_box(static_cast<size_t>(g->num_solid));
    std::vector<std::vector<oxh::Box>> prim_boxes(static_cast<size_t>(g->num_solid));
    // a prim whose root node is a sphere leaf without a transform has a known surface: the
    // sphere (centre, radius in node q0); any other prim lies somewhere in its box
    std::vector<std::vector<oxh::RadItem>> prim_rad(static_cast<size_t>(g->num_solid));
    for (int64_t s = 0; s < g->num_solid; s++) {
        const int32_t* so = g->solid + 12 * s;
        int numPrim = so[4], primOffset = so[5];
        if (numPrim < 1 || primOffset < 0 || int64_t(primOffset) + numPrim > g->num_prim)
            return bail(fail(OXR_EINVAL, "solid " + std::to_string(s) + ": prim range outside prim array"));
        std::vector<oxh::Box> boxes(static_cast<size_t>(numPrim));
        std::vector<oxh::RadItem> rads(static_cast<size_t>(numPrim));
        oxh::Box sb = oxh::Box::empty();
        for (int p = 0; p < numPrim; p++) {
            const float* q = g->prim + 16 * size_t(primOffset + p);
            oxh::Box b;
            b.lo[0] = q[8]; b.lo[1] = q[9]; b.lo[2] = q[10];
            b.hi[0] = q[11]; b.hi[1] = q[12]; b.hi[2] = q[13];
            for (int k = 0; k < 3; k++) if (!(b.lo[k] <= b.hi[k])) return bail(fail(OXR_EINVAL, "prim with empty/NaN AABB"));
            boxes[size_t(p)] = b;
            sb.grow(b);
            int32_t nodeOffset;
            std::memcpy(&nodeOffset, q + 1, 4);
            const float* nd = g->node + 16 * size_t(nodeOffset);
            uint32_t tc, gt;
            std::memcpy(&tc, nd + 14, 4);
            std::memcpy(&gt, nd + 15, 4);
            oxh::RadItem& ri = rads[size_t(p)];
            ri.sphere = tc == 101u && (gt & 0x7fffffffu) == 0u && nd[3] > 0.f;
            ri.c[0] = nd[0]; ri.c[1] = nd[1]; ri.c[2] = nd[2]; ri.r = nd[3];
        }
        prim_rad[size_t(s)] = std::move(rads);
        e->solid_prim_offset[size_t(s)] = primOffset;
        e->solid_num_prim[size_t(s)] = numPrim;
        solid_box[size_t(s)] = sb;
        prim_boxes[size_t(s)] = std::move(boxes);
    }
    // Scene extent: the largest world coordinate of any instance box (W) and the largest
    // object-frame distance a world point within W can map to (S). The device slab test
    // (reciprocal direction, fma with origin*inverse) rounds within a few float ulps of
    // max(|origin|, |box|, |t*dir|); boxes are widened by 1e-6 of that scale.
    double W = 0.0, S = 0.0;
    for (int64_t i = 0; i < g->num_inst; i++) {
        const float* q = g->inst + 16 * i;
        int32_t gas;
        std::memcpy(&gas, q + 7, 4);
        if (gas < 0 || gas >= g->num_solid) return bail(fail(OXR_EINVAL, "instance " + std::to_string(i) + ": gas_idx outside solids"));
        const oxh::Box& sb = solid_box[size_t(gas)];
        for (int c = 0; c < 8; c++) {
            double p[3] = {(c & 1) ? sb.hi[0] : sb.lo[0], (c & 2) ? sb.hi[1] : sb.lo[1], (c & 4) ? sb.hi[2] : sb.lo[2]};
            for (int k = 0; k < 3; k++) W = std::max(W, std::fabs(p[0] * q[k] + p[1] * q[4 + k] + p[2] * q[8 + k] + q[12 + k]));
        }
    }
    for (int64_t i = 0; i < g->num_inst; i++) {
        double Wi[4][3];
        invert_instance(g->inst + 16 * i, Wi);
        double lin = 0.0, tr = 0.0;
        for (int r = 0; r < 3; r++) lin = std::max(lin, std::fabs(Wi[0][r]) + std::fabs(Wi[1][r]) + std::fabs(Wi[2][r]));
        for (int k = 0; k < 3; k++) tr = std::max(tr, std::fabs(Wi[3][k]));
        if (std::isfinite(lin) && std::isfinite(tr)) S = std::max(S, 2.0 * lin * W + tr);
    }
    const double slack_world = 1e-6 * 2.0 * W, slack_obj = 1e-6 * std::max(S, 2.0 * W);
    for (int64_t s = 0; s < g->num_solid; s++) {
        oxh::Bvh bv = oxh::build_bvh(prim_boxes[size_t(s)], int(blas_all.size()), slack_obj, &prim_rad[size_t(s)]);
        e->solid_blas_root[size_t(s)] = bv.root_ref;
        max_blas_depth = std::max(max_blas_depth, bv.depth);
        blas_all.insert(blas_all.end(), bv.nodes.begin(), bv.nodes.end());
        rad_blas.insert(rad_blas.end(), bv.rad.begin(), bv.rad.end());
    }
    // ---- instances and TLAS
    std::vector<oxd::InstD> insts(static_cast<size_t>(g->num_inst));
    std::vector<oxh::Box> inst_box(static_cast<size_t>(g->num_inst));
    for (int64_t i = 0; i < g->num_inst; i++) {
        const float* q = g->inst + 16 * i;
        int32_t gas, sid1;
        std::memcpy(&gas, q + 7, 4);
        std::memcpy(&sid1, q + 11, 4);
        if (gas < 0 || gas >= g->num_solid) return bail(fail(OXR_EINVAL, "instance " + std::to_string(i) + ": gas_idx outside solids"));
        bool ident = q[0] == 1.f && q[1] == 0.f && q[2] == 0.f && q[4] == 0.f && q[5] == 1.f && q[6] == 0.f &&
                     q[8] == 0.f && q[9] == 0.f && q[10] == 1.f && q[12] == 0.f && q[13] == 0.f && q[14] == 0.f;
        double W[4][3];
        invert_instance(q, W);
        for (int r = 0; r < 4; r++) for (int k = 0; k < 3; k++)
            if (!std::isfinite(W[r][k])) return bail(fail(OXR_EINVAL, "instance transform not invertible"));
        oxd::InstD d;
        d.c0 = make_float4(float(W[0][0]), float(W[1][0]), float(W[2][0]), float(W[3][0]));
        d.c1 = make_float4(float(W[0][1]), float(W[1][1]), float(W[2][1]), float(W[3][1]));
        d.c2 = make_float4(float(W[0][2]), float(W[1][2]), float(W[2][2]), float(W[3][2]));
        uint32_t ii_id = ((uint32_t(i) & 0xffffu) << 16) | (uint32_t(sid1) & 0xffffu);
        // flags: bit 0 identity transform; bits 8-15 the instance visibility mask, one bit
        // per solid up to the eighth (Properties::visibilityMask, Properties.cc:45-55)
        const uint32_t vis = 1u << std::min<int32_t>(gas, 7);
        d.info = make_int4(e->solid_prim_offset[size_t(gas)], int(ii_id), int((ident ? 1u : 0u) | (vis << 8)),
                           e->solid_blas_root[size_t(gas)]);
        insts[size_t(i)] = d;
        // world box of the instance: solid box corners through the object->world transform
        const oxh::Box& sb = solid_box[size_t(gas)];
        oxh::Box wb = oxh::Box::empty();
        for (int c = 0; c < 8; c++) {
            double p[3] = {(c & 1) ? sb.hi[0] : sb.lo[0], (c & 2) ? sb.hi[1] : sb.lo[1], (c & 4) ? sb.hi[2] : sb.lo[2]};
            double w[3];
            for (int k = 0; k < 3; k++) w[k] = p[0] * q[k] + p[1] * q[4 + k] + p[2] * q[8 + k] + q[12 + k];
            wb.grow_point(w);
        }
        inst_box[size_t(i)] = wb;
    }
    // Flattened form, built besides the two levels while the placed prims number at most
    // kFlatMax and used when the tuning asks (bvh_mode 1): one BVH over every (instance,
    // prim) with world boxes. A trace then skips the BLAS roots and the return to the top
    // level, and prims are culled by their own world boxes rather than their instance's;
    // scenes that mix huge and small boxes (concentric shells around small instances) can
    // trace slower on it, so the autotuner picks. Each prim is still tested in its
    // instance's frame with the same arithmetic and the closest-hit key is the same, so the
    // hits are identical.
    const int64_t kFlatMax = int64_t(1) << 21;
    int64_t placed = 0;
    for (int64_t i = 0; i < g->num_inst; i++) {
        int32_t gas; std::memcpy(&gas, g->inst + 16 * i + 7, 4);
        placed += e->solid_num_prim[size_t(gas)];
    }
    e->has_flat = placed <= kFlatMax;
    oxh::Bvh tl, tf;
    std::vector<int2> flat_items;
    if (e->has_flat) {
        std::vector<oxh::Box> leaf_box;
        std::vector<oxh::RadItem> leaf_rad;
        leaf_box.reserve(size_t(placed)); flat_items.reserve(size_t(placed)); leaf_rad.reserve(size_t(placed));
        for (int64_t i = 0; i < g->num_inst; i++) {
            const float* q = g->inst + 16 * i;
            int32_t gas; std::memcpy(&gas, q + 7, 4);
            const std::vector<oxh::Box>& pb = prim_boxes[size_t(gas)];
            for (size_t p = 0; p < pb.size(); p++) {
                oxh::Box wb = oxh::Box::empty();
                for (int c = 0; c < 8; c++) {
                    double pt[3] = {(c & 1) ? pb[p].hi[0] : pb[p].lo[0], (c & 2) ? pb[p].hi[1] : pb[p].lo[1],
                                    (c & 4) ? pb[p].hi[2] : pb[p].lo[2]};
                    double w[3];
                    for (int k = 0; k < 3; k++) w[k] = pt[0] * q[k] + pt[1] * q[4 + k] + pt[2] * q[8 + k] + q[12 + k];
                    wb.grow_point(w);
                }
                leaf_box.push_back(wb);
                flat_items.push_back(make_int2(int(i), int(p)));
                // world-frame sphere only for an instance with the identity transform
                oxh::RadItem ri = prim_rad[size_t(gas)][p];
                const bool ident = q[0] == 1.f && q[1] == 0.f && q[2] == 0.f && q[4] == 0.f && q[5] == 1.f && q[6] == 0.f &&
                                   q[8] == 0.f && q[9] == 0.f && q[10] == 1.f && q[12] == 0.f && q[13] == 0.f && q[14] == 0.f;
                ri.sphere = ri.sphere && ident;
                leaf_rad.push_back(ri);
            }
        }
        tf = oxh::build_bvh(leaf_box, 0, slack_world, &leaf_rad);
        e->stack_need_f = tf.depth + 1;
        e->tlas_root_f = tf.root_ref;
        e->tlas_nodes_f = int64_t(tf.nodes.size());
    }
    std::vector<oxh::RadItem> inst_rad(inst_box.size());
    for (auto& r : inst_rad) r.sphere = false;
    tl = oxh::build_bvh(inst_box, 0, slack_world, &inst_rad);
    e->stack_need = tl.depth + 1 + max_blas_depth;
    e->tlas_root = tl.root_ref;
    e->tlas_nodes = int64_t(tl.nodes.size());
    e->blas_nodes = int64_t(blas_all.size());
    if (e->stack_cap() > 96) return bail(fail(OXR_EINVAL, "BVH too deep for the traversal stack"));

    // ---- uploads
    auto up = [&](auto& buf, const void* src, size_t count, size_t elem) -> int {
        size_t nn = std::max<size_t>(count, 1);
        hipError_t x = buf.reserve(nn);
        if (x != hipSuccess) return hip_fail(x, "hipMalloc geometry");
        if (count) {
            x = hipMemcpy(buf.p, src, count * elem, hipMemcpyHostToDevice);
            if (x != hipSuccess) return hip_fail(x, "hipMemcpy geometry");
        }
        e->geometry_bytes += int64_t(nn * elem);
        return OXR_OK;
    };
    if ((rc = up(e->node, g->node, size_t(g->num_node) * 4, 16))) return bail(rc);
    if ((rc = up(e->plan, g->plan, size_t(g->num_plan), 16))) return bail(rc);
    if ((rc = up(e->itra, g->itra, size_t(g->num_itra) * 4, 16))) return bail(rc);
    if ((rc = up(e->prim, primtab.data(), primtab.size(), 8))) return bail(rc);
    if ((rc = up(e->tlas, tl.nodes.data(), tl.nodes.size(), 64))) return bail(rc);
    if ((rc = up(e->blas, blas_all.data(), blas_all.size(), 64))) return bail(rc);
    if ((rc = up(e->flat, flat_items.data(), flat_items.size(), 8))) return bail(rc);
    if ((rc = up(e->tlas_f, tf.nodes.data(), tf.nodes.size(), 64))) return bail(rc);
    if ((rc = up(e->inst, insts.data(), insts.size(), 64))) return bail(rc);
    if ((rc = up(e->rad_tlas, tl.rad.data(), tl.rad.size() * 2, 16))) return bail(rc);
    if ((rc = up(e->rad_blas, rad_blas.data(), rad_blas.size() * 2, 16))) return bail(rc);
    if ((rc = up(e->rad_flat, tf.rad.data(), tf.rad.size() * 2, 16))) return bail(rc);
    size_t bnd_f4 = size_t(ph->num_bnd) * 4 * 2 * size_t(ph->num_wl);
    if ((rc = up(e->bnd, ph->bnd, bnd_f4, 16))) return bail(rc);
    if ((rc = up(e->optical, ph->optical, size_t(ph->num_bnd) * 4, 16))) return bail(rc);
    if (ph->icdf) { if ((rc = up(e->icdf, ph->icdf, size_t(ph->icdf_rows) * ph->icdf_width, 4))) return bail(rc); }
    e->tables.bnd = e->bnd.p;
    e->tables.optical = e->optical.p;
    e->tables.icdf = ph->icdf ? e->icdf.p : nullptr;
    e->tables.num_wl = ph->num_wl;
    e->tables.wl_min = ph->wl_min;
    e->tables.wl_step = ph->wl_step;
    e->tables.icdf_width = ph->icdf ? ph->icdf_width : 0;
    e->tables.hd_factor = ph->icdf ? ph->hd_factor : 0;
    if ((he = hipDeviceSynchronize()) != hipSuccess) return bail(hip_fail(he, "upload sync"));
    e->tune = e->effective(e->tune);
    *out = reinterpret_cast<oxr_engine*>(e);
    return OXR_OK;
}

void oxr_engine_destroy(oxr_engine* eh) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    e->ev.release_all();
    e->node.release(); e->plan.release(); e->itra.release(); e->bnd.release(); e->prim.release();
    e->tlas.release(); e->blas.release(); e->inst.release(); e->flat.release(); e->tlas_f.release();
    e->rad_tlas.release(); e->rad_blas.release(); e->rad_flat.release(); e->optical.release(); e->icdf.release();
    if (e->e0) (void)hipEventDestroy(e->e0);
    if (e->e1) (void)hipEventDestroy(e->e1);
    if (e->e2) (void)hipEventDestroy(e->e2);
    if (e->eg) (void)hipEventDestroy(e->eg);
    if (e->h_stage) (void)hipHostFree(e->h_stage);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int oxr_engine_set_config(oxr_engine* eh, const oxr_config* c) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !c) return fail(OXR_EINVAL, "null argument");
    if (!config_ok(*c)) return fail(OXR_EINVAL, "config");
    std::lock_guard<std::mutex> lk(e->mu);
    e->cfg = *c;
    if (e->cfg.max_photon == 0) e->cfg.max_photon = heuristic_max_photon(e->cfg, e->total_mem);
    e->launched = false;
    return OXR_OK;
}

int oxr_engine_set_tuning(oxr_engine* eh, const oxr_tuning* t) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!t) return fail(OXR_EINVAL, "null argument");
    if (!tuning_ok(*t)) return fail(OXR_EINVAL, "tuning out of range");
    if (!e) return fail(OXR_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(e->mu);
    e->tune = e->effective(*t);
    return OXR_OK;
}

int oxr_engine_get_tuning(oxr_engine* eh, oxr_tuning* t) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !t) return fail(OXR_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(e->mu);
    *t = e->tune;
    return OXR_OK;
}

int oxr_engine_get_config(oxr_engine* eh, oxr_config* c) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !c) return fail(OXR_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(e->mu);
    *c = e->cfg;
    return OXR_OK;
}

int oxr_engine_query(oxr_engine* eh, oxr_engine_info* info) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !info) return fail(OXR_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(e->mu);   // tuning and event buffers are read below
    std::memset(info, 0, sizeof(*info));
    info->num_solid = e->num_solid; info->num_prim = e->num_prim; info->num_node = e->num_node; info->num_inst = e->num_inst;
    info->tlas_nodes = e->tlas_nodes; info->blas_nodes = e->blas_nodes;
    info->node_bytes = int32_t(sizeof(oxd::BNode));
    info->num_cu = e->num_cu;
    info->device_bytes = e->geometry_bytes;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) { info->free_bytes = int64_t(fr); info->total_bytes = int64_t(tot); }
    std::strncpy(info->device_name, e->name.c_str(), sizeof(info->device_name) - 1);
    std::strncpy(info->arch, e->arch.c_str(), sizeof(info->arch) - 1);
    info->csg_height = e->csg_height;
    info->csg_in_registers = e->tune.csg_stack == 1 && e->registers_csg_ok() ? 1 : 0;
    info->simple_leaves = e->simple_leaves && !e->has_list ? 1 : 0;
    info->csg_in_lds = info->simple_leaves && !e->cfg.fp64_roots &&
                       e->params(e->ev, 0, 0, 0, e->tune.block_size).csg_node4 > 0 ? 1 : 0;
    return OXR_OK;
}

// The engine's current event (QEvt role); callers hold e.mu.
static int set_gensteps_locked(Engine& e, const float* gs, int64_t num_gs, int32_t event_index,
                               uint64_t photon_slot_offset, int64_t* num_photon) {
    int64_t n = 0;
    int rc = upload_gensteps(e, e.ev, e.stream, gs, num_gs, &n);
    if (rc) return rc;
    e.h_gs.assign(gs, gs + 24 * num_gs);
    e.num_gs = num_gs;
    e.num_photon = n;
    e.event_index = event_index;
    e.slot_offset = photon_slot_offset;
    e.launched = false;
    *num_photon = n;
    return OXR_OK;
}

int oxr_event_set_gensteps(oxr_engine* eh, const float* gs, int64_t num_gs, int32_t event_index,
                           uint64_t photon_slot_offset, int64_t* num_photon) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !num_photon) return fail(OXR_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(e->mu);
    return set_gensteps_locked(*e, gs, num_gs, event_index, photon_slot_offset, num_photon);
}

int oxr_event_set_input_photons(oxr_engine* eh, const float* photons, int64_t n) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || n < 0 || (n > 0 && !photons)) return fail(OXR_EINVAL, "bad input photons");
    std::lock_guard<std::mutex> lk(e->mu);
    if (n == 0) { e->have_input = false; e->input_n = 0; return OXR_OK; }
    OXR_HIP(e->ev.input.reserve(size_t(n) * 4));
    OXR_HIP(hipMemcpy(e->ev.input.p, photons, size_t(n) * 64, hipMemcpyHostToDevice));
    e->have_input = true;
    e->input_n = n;
    return OXR_OK;
}

int oxr_event_set_precooked(oxr_engine* eh, const float* u, int64_t n_photon, int32_t n_per) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return fail(OXR_EINVAL, "null engine");
    std::lock_guard<std::mutex> lk(e->mu);
    if (!u || n_photon == 0) { e->have_pre = false; e->npre = 0; return OXR_OK; }
    if (n_photon < 0 || n_per < 1) return fail(OXR_EINVAL, "precooked shape");
    OXR_HIP(e->ev.pre.reserve(size_t(n_photon) * size_t(n_per)));
    OXR_HIP(hipMemcpy(e->ev.pre.p, u, size_t(n_photon) * size_t(n_per) * 4, hipMemcpyHostToDevice));
    e->have_pre = true;
    e->npre = n_per;
    e->pre_rows = n_photon;
    return OXR_OK;
}

static int check_sources(Engine& e, int64_t n, uint64_t slot_off) {
    // photons of INPUT_PHOTON gensteps read input[slot_offset + idx] (qsim.h:2537 reads the
    // input array at the absolute photon index); precooked rows read pre[idx]
    int64_t need = 0, off = 0;
    for (int64_t g = 0; g < e.num_gs; g++) {
        off += int64_t(fbits(e.h_gs[size_t(24 * g + 3)]));
        if (fbits(e.h_gs[size_t(24 * g)]) == 19u) need = int64_t(slot_off) + off;
    }
    if (need > 0 && (!e.have_input || need > e.input_n))
        return fail(OXR_ESTATE, "INPUT_PHOTON gensteps need oxr_event_set_input_photons covering their photon indices");
    if (e.have_pre && e.pre_rows < n)
        return fail(OXR_ESTATE, "precooked stream has fewer rows than photons");
    return OXR_OK;
}

// One launch over the current event; callers hold e.mu. The event buffers are sized
// again for the current config first: oxr_engine_set_config may have switched on or
// enlarged a debug array (seq, record, rec, prd, photonlite) after the gensteps came.
static int simulate_launch_locked(Engine& e, int32_t form, bool count, oxr_launch_stats* st) {
    if (!count && e.num_gs == 0 && e.num_photon == 0) {
        if (st) std::memset(st, 0, sizeof(*st));
        e.launched = true;
        e.num_hit = 0;
        return OXR_OK;
    }
    int rc = check_sources(e, e.num_photon, e.slot_offset);
    if (rc) return rc;
    rc = ensure_event_capacity(e, e.ev, e.num_photon);
    if (rc) return rc;
    rc = run_launch(e, e.ev, e.stream, e.num_photon, int64_t(e.slot_offset), e.event_index, form, count, st,
                    e.have_pre ? e.ev.pre.p : nullptr, e.have_pre ? e.npre : 0,
                    e.have_input ? e.ev.input.p : nullptr);
    if (rc == OXR_OK) e.launched = true;
    return rc;
}

int oxr_simulate_launch(oxr_engine* eh, int32_t form, oxr_launch_stats* st) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return fail(OXR_EINVAL, "null engine");
    if (form != OXR_FORM_PLAIN && form != OXR_FORM_PERSISTENT) return fail(OXR_EINVAL, "unknown form");
    std::lock_guard<std::mutex> lk(e->mu);
    return simulate_launch_locked(*e, form, false, st);
}

// Counting launch (node visits / prim tests) on the current event: test and
// measurement helper, not a separate form.
int oxr_simulate_launch_counting(oxr_engine* eh, oxr_launch_stats* st) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e) return fail(OXR_EINVAL, "null engine");
    std::lock_guard<std::mutex> lk(e->mu);
    return simulate_launch_locked(*e, OXR_FORM_PLAIN, true, st);
}

int oxr_event_get(oxr_engine* eh, int32_t what, void* dst, int64_t cap, int64_t* count) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !count) return fail(OXR_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(e->mu);
    const void* src = nullptr;
    int64_t n = 0;
    size_t item = 0;
    switch (what) {
        case OXR_GET_PHOTON: src = e->ev.photon.p; n = e->num_photon; item = 64; break;
        case OXR_GET_HIT: src = e->ev.hit.p; n = e->num_hit; item = 64; break;
        case OXR_GET_SEQ:
            if (!e->cfg.seq_enabled) return fail(OXR_ESTATE, "seq_enabled is off");
            src = e->ev.seq.p; n = e->num_photon; item = 32; break;
        case OXR_GET_RECORD:
            if (e->cfg.max_record <= 0) return fail(OXR_ESTATE, "max_record is 0");
            src = e->ev.record.p; n = e->num_photon * e->cfg.max_record; item = 64; break;
        case OXR_GET_REC:
            if (e->cfg.max_rec <= 0) return fail(OXR_ESTATE, "max_rec is 0");
            src = e->ev.rec.p; n = e->num_photon * e->cfg.max_rec; item = 16; break;
        case OXR_GET_PRD:
            if (e->cfg.max_prd <= 0) return fail(OXR_ESTATE, "max_prd is 0");
            src = e->ev.prd.p; n = e->num_photon * e->cfg.max_prd; item = 32; break;
        case OXR_GET_SEED: src = e->ev.seed.p; n = e->num_photon; item = 4; break;
        case OXR_GET_GENSTEP: src = e->ev.gs.p; n = e->num_gs; item = 96; break;
        case OXR_GET_PHOTONLITE:
            if (!e->cfg.photonlite) return fail(OXR_ESTATE, "photonlite is off");
            src = e->ev.lite.p; n = e->num_photon; item = 16; break;
        case OXR_GET_HITLITE:
            if (!e->cfg.photonlite) return fail(OXR_ESTATE, "photonlite is off");
            if (e->launched && !e->hitlite_ready) {
                // same selection and order as the sphoton hits (sphotonlite_selector,
                // sphotonlite.h:159-165): reuse the launch's compaction offsets
                OXR_HIP(e->ev.hitlite.reserve(size_t(std::max<int64_t>(e->num_hit, 1))));
                OXR_HIP(oxl::scatter_lite(e->stream, e->ev.hitflag.p, e->ev.lite.p, e->num_photon, e->ev.offs.p,
                                          e->ev.hitlite.p, e->num_hit));
                OXR_HIP(hipStreamSynchronize(e->stream));
                e->hitlite_ready = true;
            }
            src = e->ev.hitlite.p; n = e->num_hit; item = 16; break;
        default: return fail(OXR_EINVAL, "unknown gather kind");
    }
    if (what != OXR_GET_SEED && what != OXR_GET_GENSTEP && !e->launched) return fail(OXR_ESTATE, "no launch for this event");
    *count = n;
    int64_t m = std::min(n, std::max<int64_t>(cap, 0));
    if (m > 0) {
        if (!dst) return fail(OXR_EINVAL, "null destination");
        int rc = e->d2h(dst, src, size_t(m) * item);
        if (rc) return rc;
    }
    return m < n ? fail(OXR_ECAPACITY, "destination smaller than count") : OXR_OK;
}

// Device-to-host copy through the engine's pinned buffer, in chunks of at most 8 MB on the
// engine stream (after the work that produced src): a copy straight into pageable memory
// goes through the runtime's own staging at a few GB/s.
int Engine::d2h(void* dst, const void* src, size_t bytes) {
    constexpr size_t kStage = size_t(8) << 20;
    if (bytes == 0) return OXR_OK;
    const size_t want = std::min(bytes, kStage);
    if (h_stage_bytes < want) {
        if (h_stage) (void)hipHostFree(h_stage);
        h_stage = nullptr;
        h_stage_bytes = 0;
        OXR_HIP(hipHostMalloc(&h_stage, want, hipHostMallocDefault));
        h_stage_bytes = want;
    }
    for (size_t off = 0; off < bytes; off += h_stage_bytes) {
        const size_t k = std::min(h_stage_bytes, bytes - off);
        OXR_HIP(hipMemcpyAsync(h_stage, static_cast<const char*>(src) + off, k, hipMemcpyDeviceToHost, stream));
        OXR_HIP(hipStreamSynchronize(stream));
        std::memcpy(static_cast<char*>(dst) + off, h_stage, k);
    }
    return OXR_OK;
}

static int copy_out(Engine& e, const void* src, int64_t n, size_t item, void* dst, int64_t cap, int64_t* count) {
    *count = n;
    int64_t m = std::min(n, std::max<int64_t>(cap, 0));
    if (m > 0) {
        if (!dst) return fail(OXR_EINVAL, "null destination");
        int rc = e.d2h(dst, src, size_t(m) * item);
        if (rc) return rc;
    }
    return m < n ? fail(OXR_ECAPACITY, "destination smaller than count") : OXR_OK;
}

int oxr_event_get_merged(oxr_engine* eh, int32_t what, float time_window, void* dst, int64_t cap, int64_t* count) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !count) return fail(OXR_EINVAL, "null argument");
    if (what != OXR_GET_HIT && what != OXR_GET_HITLITE) return fail(OXR_EINVAL, "merge needs OXR_GET_HIT or OXR_GET_HITLITE");
    if (!(time_window >= 0.f)) return fail(OXR_EINVAL, "time_window must be >= 0");
    std::lock_guard<std::mutex> lk(e->mu);
    if (!e->launched) return fail(OXR_ESTATE, "no launch for this event");
    bool lite = what == OXR_GET_HITLITE;
    if (lite && !e->cfg.photonlite) return fail(OXR_ESTATE, "photonlite is off");
    size_t words = lite ? 4 : 16;
    const void* in = lite ? static_cast<const void*>(e->ev.lite.p) : static_cast<const void*>(e->ev.photon.p);
    OXR_HIP(e->ev.merged.reserve(words * size_t(std::max<int64_t>(e->num_photon, 1))));
    long long nout = 0;
    OXR_HIP(oxl::merge(e->stream, lite, in, e->num_photon, e->cfg.hitmask, true, time_window, e->ev.merged.p, &nout));
    return copy_out(*e, e->ev.merged.p, nout, words * 4, dst, cap, count);
}

int oxr_merge_hits(oxr_engine* eh, int32_t what, float time_window, const void* in, int64_t n, void* dst, int64_t cap,
                   int64_t* count) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !count || n < 0 || (n > 0 && !in)) return fail(OXR_EINVAL, "bad argument");
    if (what != OXR_GET_HIT && what != OXR_GET_HITLITE) return fail(OXR_EINVAL, "merge needs OXR_GET_HIT or OXR_GET_HITLITE");
    if (!(time_window >= 0.f)) return fail(OXR_EINVAL, "time_window must be >= 0");
    std::lock_guard<std::mutex> lk(e->mu);
    bool lite = what == OXR_GET_HITLITE;
    size_t words = lite ? 4 : 16;
    if (n == 0) { *count = 0; return OXR_OK; }
    DBuf<uint32_t> din, dout;
    auto done = [&](int rc) { din.release(); dout.release(); return rc; };
    hipError_t he;
    if ((he = din.reserve(words * size_t(n))) != hipSuccess) return done(hip_fail(he, "merge input"));
    if ((he = dout.reserve(words * size_t(n))) != hipSuccess) return done(hip_fail(he, "merge output"));
    if ((he = hipMemcpy(din.p, in, words * 4 * size_t(n), hipMemcpyHostToDevice)) != hipSuccess) return done(hip_fail(he, "merge upload"));
    long long nout = 0;
    if ((he = oxl::merge(e->stream, lite, din.p, n, 0u, false, time_window, dout.p, &nout)) != hipSuccess)
        return done(hip_fail(he, "merge"));
    return done(copy_out(*e, dout.p, nout, words * 4, dst, cap, count));
}

int oxr_simulate(oxr_engine* eh, const float* gs, int64_t num_gs, int32_t event_index, int32_t form, void* hits,
                 int64_t hit_cap, int64_t* num_hit, oxr_launch_stats* st) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !num_hit || num_gs < 0 || (num_gs > 0 && !gs)) return fail(OXR_EINVAL, "bad argument");
    if (form != OXR_FORM_PLAIN && form != OXR_FORM_PERSISTENT) return fail(OXR_EINVAL, "unknown form");
    oxr_launch_stats acc{};
    int64_t total_hit = 0;
    uint64_t slot = 0;
    int64_t g0 = 0;
    bool overflow = false;
    // one lock over every slice's upload, launch and gather: no other call on this engine
    // can replace the event between a launch and the copy of its hits
    std::lock_guard<std::mutex> lk(e->mu);
    while (g0 < num_gs || (num_gs == 0 && g0 == 0)) {
        // genstep slice bounded by max_photon (SGenstep::GetGenstepSlices role)
        int64_t g1 = g0, np = 0;
        while (g1 < num_gs) {
            int64_t k = int64_t(fbits(gs[24 * g1 + 3]));
            if (k > e->cfg.max_photon) return fail(OXR_EINVAL, "single genstep exceeds max_photon");
            if (np + k > e->cfg.max_photon) break;
            np += k; g1++;
        }
        int64_t n = 0;
        int rc = set_gensteps_locked(*e, gs + 24 * g0, g1 - g0, event_index, slot, &n);
        if (rc) return rc;
        oxr_launch_stats s{};
        rc = simulate_launch_locked(*e, form, false, &s);
        if (rc) return rc;
        int64_t room = std::max<int64_t>(0, hit_cap - total_hit);
        int64_t take = std::min(room, s.num_hit);
        if (take > 0 && hits) {
            int rcc = e->d2h(static_cast<char*>(hits) + size_t(total_hit) * 64, e->ev.hit.p, size_t(take) * 64);
            if (rcc) return rcc;
        }
        if (take < s.num_hit) overflow = true;
        total_hit += s.num_hit;
        acc.num_photon += s.num_photon; acc.num_hit += s.num_hit; acc.bounce_steps += s.bounce_steps;
        acc.kernel_ms += s.kernel_ms; acc.compact_ms += s.compact_ms; acc.total_ms += s.total_ms;
        acc.generate_ms += s.generate_ms;
        slot += uint64_t(n);
        if (num_gs == 0) break;
        g0 = g1;
    }
    *num_hit = total_hit;
    if (st) *st = acc;
    return overflow ? fail(OXR_ECAPACITY, "hit buffer smaller than hit count") : OXR_OK;
}

int oxr_trace_rays(oxr_engine* eh, const float* rays, int64_t n, float* prd) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || n < 0 || (n > 0 && (!rays || !prd))) return fail(OXR_EINVAL, "bad argument");
    if (n == 0) return OXR_OK;
    std::lock_guard<std::mutex> lk(e->mu);
    DBuf<float4> dr, dp;
    auto cleanup = [&]() { dr.release(); dp.release(); };
    hipError_t x;
    if ((x = dr.reserve(size_t(n) * 2)) != hipSuccess) { cleanup(); return hip_fail(x, "hipMalloc rays"); }
    if ((x = dp.reserve(size_t(n) * 2)) != hipSuccess) { cleanup(); return hip_fail(x, "hipMalloc prd"); }
    if ((x = hipMemcpy(dr.p, rays, size_t(n) * 32, hipMemcpyHostToDevice)) != hipSuccess) { cleanup(); return hip_fail(x, "copy rays"); }
    EventBuffers dummy;
    oxd::Params P = e->params(dummy, 0, 0, 0, 256);
    if ((x = oxl::trace_rays(e->stream, P, dr.p, n, dp.p, e->cfg.fp64_roots != 0, e->has_list)) != hipSuccess) { cleanup(); return hip_fail(x, "trace launch"); }
    if ((x = hipStreamSynchronize(e->stream)) != hipSuccess) { cleanup(); return hip_fail(x, "trace sync"); }
    x = hipMemcpy(prd, dp.p, size_t(n) * 32, hipMemcpyDeviceToHost);
    cleanup();
    if (x != hipSuccess) return hip_fail(x, "copy prd");
    return OXR_OK;
}

int oxr_simtrace(oxr_engine* eh, const float* gs, int64_t num_gs, float* simtrace, int64_t capacity, int64_t* num_out) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !num_out || num_gs < 0 || capacity < 0 || (num_gs > 0 && !gs)) return fail(OXR_EINVAL, "bad argument");
    *num_out = 0;
    std::vector<long long> off(size_t(num_gs) + 1, 0);
    for (int64_t g = 0; g < num_gs; g++) {
        const float* q = gs + 24 * g;
        const uint32_t code = fbits(q[0]), axes = fbits(q[1]);
        const int32_t np = int32_t(fbits(q[3]));
        if (code != 17) return fail(OXR_EINVAL, "simtrace takes FRAME gensteps (gencode 17) only");
        if (axes > 3u) return fail(OXR_EINVAL, "FRAME genstep grid axes must be XYZ, YZ, XZ or XY (0-3)");
        if (np < 0) return fail(OXR_EINVAL, "FRAME genstep with a negative photon count");
        off[size_t(g) + 1] = off[size_t(g)] + np;
    }
    const int64_t n = off[size_t(num_gs)];
    *num_out = n;
    if (n > capacity) return fail(OXR_ECAPACITY, "simtrace buffer too small");
    if (n == 0) return OXR_OK;
    if (!simtrace) return fail(OXR_EINVAL, "null simtrace buffer");
    if (num_gs > 0x7fffffffLL) return fail(OXR_EINVAL, "too many gensteps");
    std::lock_guard<std::mutex> lk(e->mu);
    DBuf<float4> dg, dout;
    DBuf<long long> doff;
    auto cleanup = [&]() { dg.release(); dout.release(); doff.release(); };
    hipError_t x;
    if ((x = dg.reserve(size_t(num_gs) * 6)) != hipSuccess || (x = doff.reserve(size_t(num_gs) + 1)) != hipSuccess ||
        (x = dout.reserve(size_t(n) * 4)) != hipSuccess) {
        cleanup(); return hip_fail(x, "hipMalloc simtrace");
    }
    if ((x = hipMemcpy(dg.p, gs, size_t(num_gs) * 96, hipMemcpyHostToDevice)) != hipSuccess ||
        (x = hipMemcpy(doff.p, off.data(), off.size() * sizeof(long long), hipMemcpyHostToDevice)) != hipSuccess) {
        cleanup(); return hip_fail(x, "copy FRAME gensteps");
    }
    EventBuffers dummy;
    oxd::Params P = e->params(dummy, 0, 0, 0, 256);
    if ((x = oxl::simtrace(e->stream, P, dg.p, doff.p, int(num_gs), n, dout.p, e->cfg.fp64_roots != 0, e->has_list)) != hipSuccess) {
        cleanup(); return hip_fail(x, "simtrace launch");
    }
    if ((x = hipStreamSynchronize(e->stream)) != hipSuccess) { cleanup(); return hip_fail(x, "simtrace sync"); }
    x = hipMemcpy(simtrace, dout.p, size_t(n) * 64, hipMemcpyDeviceToHost);
    cleanup();
    if (x != hipSuccess) return hip_fail(x, "copy simtrace");
    return OXR_OK;
}

int oxr_intersect_prim(oxr_engine* eh, int32_t prim_index, const float* rays, int64_t n, float* isect, int32_t* valid) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || n < 0 || (n > 0 && (!rays || !isect || !valid))) return fail(OXR_EINVAL, "bad argument");
    if (prim_index < 0 || prim_index >= e->num_prim) return fail(OXR_EINVAL, "prim index out of range");
    if (n == 0) return OXR_OK;
    std::lock_guard<std::mutex> lk(e->mu);
    int2 pr;
    OXR_HIP(hipMemcpy(&pr, e->prim.p + prim_index, sizeof(int2), hipMemcpyDeviceToHost));
    DBuf<float4> dr, di;
    DBuf<int> dv;
    auto cleanup = [&]() { dr.release(); di.release(); dv.release(); };
    hipError_t x;
    if ((x = dr.reserve(size_t(n) * 2)) != hipSuccess || (x = di.reserve(size_t(n))) != hipSuccess || (x = dv.reserve(size_t(n))) != hipSuccess) {
        cleanup(); return hip_fail(x, "hipMalloc isect");
    }
    if ((x = hipMemcpy(dr.p, rays, size_t(n) * 32, hipMemcpyHostToDevice)) != hipSuccess) { cleanup(); return hip_fail(x, "copy rays"); }
    if ((x = oxl::intersect_prim(e->stream, e->geom(), pr.x, dr.p, n, di.p, dv.p, e->cfg.fp64_roots != 0, e->has_list)) != hipSuccess) { cleanup(); return hip_fail(x, "isect launch"); }
    if ((x = hipStreamSynchronize(e->stream)) != hipSuccess) { cleanup(); return hip_fail(x, "isect sync"); }
    if ((x = hipMemcpy(isect, di.p, size_t(n) * 16, hipMemcpyDeviceToHost)) != hipSuccess) { cleanup(); return hip_fail(x, "copy isect"); }
    x = hipMemcpy(valid, dv.p, size_t(n) * 4, hipMemcpyDeviceToHost);
    cleanup();
    if (x != hipSuccess) return hip_fail(x, "copy valid");
    return OXR_OK;
}

int oxr_autotune(oxr_engine* eh, int32_t repeat, double* results, int32_t max_results, int32_t* num_results) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || repeat < 1 || max_results < 0) return fail(OXR_EINVAL, "bad argument");
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->num_photon <= 0) return fail(OXR_ESTATE, "autotune needs an event with photons");
    oxr_tuning base = e->tune, best = e->tune;
    double best_ms = 1e300;
    int nres = 0;
    // every row records the tuning that ran (Engine::effective); a candidate that runs the
    // same as one already measured is not launched again
    std::vector<oxr_tuning> seen;
    std::vector<double> seen_ms;     // the time each row of seen reports
    std::vector<char> final_ms;      // 1: that time is the mean of the head-to-head rounds
    auto same = [](const oxr_tuning& a, const oxr_tuning& b) { return std::memcmp(&a, &b, sizeof(a)) == 0; };
    auto time_launches = [&](const oxr_tuning& t, double& ms) -> int {
        e->tune = t;
        double tot = 0.0;
        for (int r = 0; r < repeat; r++) {
            oxr_launch_stats s{};
            int rc = simulate_launch_locked(*e, OXR_FORM_PERSISTENT, false, &s);
            if (rc) return rc;
            tot += s.kernel_ms;
        }
        ms = tot / repeat;
        return OXR_OK;
    };
    auto measure = [&](const oxr_tuning& t_in, double& ms) -> int {
        if (!tuning_ok(t_in)) return fail(OXR_EINVAL, "tuning out of range");
        const oxr_tuning t = e->effective(t_in);
        for (const oxr_tuning& x : seen) if (same(x, t)) { ms = 1e300; return OXR_OK; }
        int rc = time_launches(t, ms);
        if (rc) return rc;
        seen.push_back(t); seen_ms.push_back(ms); final_ms.push_back(0);
        if (results && nres < max_results) {
            double* r = results + OXR_AUTOTUNE_COLUMNS * nres;
            r[0] = t.block_size; r[1] = t.blocks_per_cu; r[2] = t.lds_top_nodes; r[3] = t.refill_batch;
            r[4] = t.stack_mode; r[5] = t.bvh_mode; r[6] = t.csg_stack; r[7] = t.bvh_radial;
            r[8] = t.photon_order; r[9] = t.generate_inplace; r[10] = ms;
        }
        nres++;
        if (ms < best_ms) { best_ms = ms; best = t; }
        return OXR_OK;
    };
    // coordinate search, two passes (the second from the first's best, so a BVH form that
    // loses at the starting launch shape is tried again at the best one): BVH form and
    // radial culling, launch shape, LDS staging, refill granularity, the traversal stack
    // form, the CSG stack depth, the photon order, then where photons are generated.
    // Tunings already measured are not launched again.
    auto search = [&](const oxr_tuning& start) -> int {
        for (int bm : {0, 1}) {
            for (int br : {0, 1}) {
                if (bm == 1 && !e->has_flat) continue;
                oxr_tuning t = start; t.bvh_mode = bm; t.bvh_radial = br;
                double ms; int rc = measure(t, ms); if (rc) return rc;
            }
        }
        const oxr_tuning b0 = best;
        for (int bs : {64, 128, 256}) {
            for (int bpc : {1, 2, 3, 4, 6, 8}) {
                oxr_tuning t = b0; t.block_size = bs; t.blocks_per_cu = bpc;
                if (oxl::simulate_smem(e->params(e->ev, 0, 0, 0, bs), bs) > 64 * 1024) continue;
                double ms; int rc = measure(t, ms); if (rc) return rc;
            }
        }
        const oxr_tuning b1 = best;
        for (int nt : {0, 16, 64, 256, 512}) {
            oxr_tuning t = b1; t.lds_top_nodes = nt;
            double ms; int rc = measure(t, ms); if (rc) return rc;
        }
        const oxr_tuning b2 = best;
        for (int rb : {1, 8, 16, 32, 48, 64}) {
            oxr_tuning t = b2; t.refill_batch = rb;
            double ms; int rc = measure(t, ms); if (rc) return rc;
        }
        const oxr_tuning b3 = best;
        for (int sm : {0, 1}) {
            oxr_tuning t = b3; t.stack_mode = sm;
            double ms; int rc = measure(t, ms); if (rc) return rc;
        }
        if (e->registers_csg_ok()) {
            const oxr_tuning b4 = best;
            for (int cs : {0, 1}) {
                oxr_tuning t = b4; t.csg_stack = cs;
                double ms; int rc = measure(t, ms); if (rc) return rc;
            }
        }
        const oxr_tuning b5 = best;
        for (int po : {0, 1}) {
            oxr_tuning t = b5; t.photon_order = po;
            double ms; int rc = measure(t, ms); if (rc) return rc;
        }
        const oxr_tuning b6 = best;
        for (int gi : {0, 1}) {
            oxr_tuning t = b6; t.generate_inplace = gi;
            double ms; int rc = measure(t, ms); if (rc) return rc;
        }
        return OXR_OK;
    };
    for (int pass = 0; pass < 2; pass++) {
        int rc = search(pass == 0 ? base : best);
        if (rc) { e->tune = base; return rc; }
    }
    // The least of many single timings is biased low (a noisy row can win by luck), so the
    // four fastest rows are timed again in turn over three rounds, one untimed launch before
    // each, and report the mean of those rounds; while the fastest row is one not yet timed
    // that way, it is timed so too. The engine is left on the fastest row.
    auto head_to_head = [&](const std::vector<size_t>& rows) -> int {
        std::vector<double> sum(rows.size(), 0.0);
        for (int round = 0; round < 3; round++) {
            for (size_t k = 0; k < rows.size(); k++) {
                double ms;
                int rc = time_launches(seen[rows[k]], ms);   // untimed launch
                if (!rc) rc = time_launches(seen[rows[k]], ms);
                if (rc) return rc;
                sum[k] += ms;
            }
        }
        for (size_t k = 0; k < rows.size(); k++) {
            const size_t i = rows[k];
            seen_ms[i] = sum[k] / 3.0;
            final_ms[i] = 1;
            if (results && int(i) < max_results) results[OXR_AUTOTUNE_COLUMNS * i + 10] = seen_ms[i];
        }
        return OXR_OK;
    };
    std::vector<size_t> order(seen.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return seen_ms[a] < seen_ms[b]; });
    order.resize(std::min<size_t>(order.size(), 4));
    for (int rc = head_to_head(order); ; rc = head_to_head(order)) {
        if (rc) { e->tune = base; return rc; }
        size_t fastest = 0;
        for (size_t i = 1; i < seen.size(); i++) if (seen_ms[i] < seen_ms[fastest]) fastest = i;
        best = seen[fastest];
        if (final_ms[fastest]) break;
        order.assign(1, fastest);
    }
    e->tune = best;
    if (num_results) *num_results = std::min(nres, max_results);   // rows written to results
    return OXR_OK;
}

} // extern "C"
