// oxr_server.cpp -- serving form: a request loop that keeps geometry, tables and the RNG
// definition resident and batches genstep requests from many events into shared device
// steps (continuous batching).
//
//  * Admission: requests enter in arrival order while the active-slot count and the HBM
//    budget (64 B/photon state + 4 B seed + reserved hit bytes + gensteps) allow.
//  * Steps: every step gives the active requests a share of the step's photon budget,
//    shortest remaining first; one persistent (or plain) launch covers all of them via
//    a segment table; segments start on compaction-chunk boundaries so each request's
//    hits are a contiguous range of the ordered hit array.
//  * Overlap: two step slots on two streams. While the device runs step k, the host
//    finishes step k-1: each segment's hits are copied device to device into blocks
//    owned by its request, and a request completes as soon as the last step holding
//    its photons has been finished. Hits cross PCIe once, in take_hits, straight into
//    the caller's buffer on the caller's thread.
//  * No allocation or free on the request path: a request's gensteps and genstep
//    offsets live in blocks of a server-owned device pool (a hipFree would synchronise
//    the device and stall the overlapped steps), and served photons find their genstep
//    by binary search of the offsets instead of a per-photon seed array.
// Photon random streams are keyed by (event_index, photon index within the event), so a
// request's result does not depend on what it was batched with.
#include "oxr_engine.h"
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>

namespace oxh {

int genstep_offsets(Engine& e, const float* gs, int64_t num_gs, std::vector<long long>& off, int64_t* n);

// Device blocks in power-of-two size classes, kept for reuse until the server ends.
// Used by the serving thread and by take_hits on the caller's thread.
struct Pool {
    std::mutex m;
    std::multimap<size_t, void*> free_;
    std::vector<void*> all;
    ~Pool() { for (void* p : all) (void)hipFree(p); }
    static size_t cls(size_t bytes) { size_t c = 256; while (c < bytes) c <<= 1; return c; }
    hipError_t get(size_t bytes, void** p, size_t* cap) {
        size_t c = cls(bytes);
        std::lock_guard<std::mutex> lk(m);
        auto it = free_.find(c);
        if (it != free_.end()) { *p = it->second; *cap = c; free_.erase(it); return hipSuccess; }
        hipError_t e = hipMalloc(p, c);
        if (e == hipSuccess) { all.push_back(*p); *cap = c; }
        return e;
    }
    void put(void* p, size_t cap) {
        if (!p) return;
        std::lock_guard<std::mutex> lk(m);
        free_.emplace(cap, p);
    }
};

struct HitBlock { void* p; size_t cap; int64_t n; };

struct Req {
    int64_t id = 0;
    int32_t event_index = 0;
    std::vector<float> gs;
    int64_t num_gs = 0, num_photon = 0;
    std::vector<long long> off;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int status = 0;
    int64_t bytes = 0;
    void* d_gs = nullptr;     // pool block: gensteps, then the num_gs+1 offsets
    size_t d_cap = 0;
    int64_t dispatched = 0;   // photons placed in launched steps
    int inflight = 0;         // launched steps holding its photons, not yet distributed
    std::vector<HitBlock> hits;   // device blocks, in photon order
    int64_t nhit = 0;
    const float4* gs_dev() const { return static_cast<const float4*>(d_gs); }
    const long long* off_dev() const {
        return reinterpret_cast<const long long*>(static_cast<const char*>(d_gs) + size_t(num_gs) * 96);
    }
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    EventBuffers buf;
    DBuf<oxd::Seg> d_segs;
    std::vector<oxd::Seg> segs;
    std::vector<Req*> owner;
    std::vector<int64_t> take;
    long long* h_offs = nullptr;    // pinned
    size_t h_offs_n = 0;
    int64_t N = 0;
    bool busy = false;
    uint64_t seqno = 0;
    std::chrono::steady_clock::time_point t0;
};

struct Server {
    Engine* e = nullptr;
    oxr_server_config cfg{};
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::unique_ptr<Req>> incoming;
    std::vector<std::unique_ptr<Req>> active;
    std::deque<int64_t> done_ids;
    std::map<int64_t, std::unique_ptr<Req>> finished;
    int64_t budget = 0, used = 0;
    bool stop = false;
    Slot slot[2];
    uint64_t next_seq = 0;
    hipStream_t admit_stream = nullptr;
    Pool pool;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double stat_steps = 0, stat_photons = 0, stat_busy_ms = 0, stat_kernel_ms = 0, stat_bounce = 0, stat_hits = 0;
    double clock() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
    void loop();
    int admit_upload(Req& r);
    void release(Req& r) { pool.put(r.d_gs, r.d_cap); r.d_gs = nullptr; r.d_cap = 0; }
    void release_hits(Req& r) { for (auto& b : r.hits) pool.put(b.p, b.cap); r.hits.clear(); }
    int launch_step(Slot& s);
    int finalize_step(Slot& s);
    void complete(std::unique_ptr<Req>& r, int status);
    bool dispatchable() const {
        for (auto& r : active) if (r->dispatched < r->num_photon) return true;
        return false;
    }
};

// queue the request's gensteps + offsets for upload on the admission stream (the loop
// synchronises that stream once per turn, before any step can use them)
int Server::admit_upload(Req& r) {
    if (r.num_gs <= 0) return OXR_OK;
    size_t gbytes = size_t(r.num_gs) * 96, obytes = (size_t(r.num_gs) + 1) * sizeof(long long);
    OXR_HIP(pool.get(gbytes + obytes, &r.d_gs, &r.d_cap));
    // one host staging copy so a single transfer carries both arrays
    std::vector<char> h(gbytes + obytes);
    std::memcpy(h.data(), r.gs.data(), gbytes);
    std::memcpy(h.data() + gbytes, r.off.data(), obytes);
    OXR_HIP(hipMemcpyAsync(r.d_gs, h.data(), h.size(), hipMemcpyHostToDevice, admit_stream));
    OXR_HIP(hipStreamSynchronize(admit_stream));   // h is pageable and local
    return OXR_OK;
}

void Server::complete(std::unique_ptr<Req>& r, int status) {
    r->status = status;
    r->t_done = clock();
    release(*r);
    std::lock_guard<std::mutex> lk(mu);
    used -= r->bytes;
    done_ids.push_back(r->id);
    finished[r->id] = std::move(r);
    cv_done.notify_all();
}

int Server::launch_step(Slot& s) {
    const long long CH = 4096;
    std::vector<Req*> order;
    for (auto& r : active) if (r->dispatched < r->num_photon) order.push_back(r.get());
    std::sort(order.begin(), order.end(), [](const Req* a, const Req* b) {
        return (a->num_photon - a->dispatched) < (b->num_photon - b->dispatched);
    });
    s.segs.clear(); s.owner.clear(); s.take.clear();
    long long cursor = 0;
    int64_t left = std::max<int64_t>(cfg.max_step_photons, CH);
    for (Req* r : order) {
        if (left <= 0) break;
        int64_t t = std::min(left, r->num_photon - r->dispatched);
        oxd::Seg g{};
        g.start = cursor; g.count = t;
        g.photon0 = uint64_t(r->dispatched);
        g.gs = r->gs_dev(); g.off = r->off_dev(); g.num_gs = int(r->num_gs);
        g.event_index = r->event_index;
        s.segs.push_back(g); s.owner.push_back(r); s.take.push_back(t);
        cursor += ((t + CH - 1) / CH) * CH;
        left -= t;
    }
    if (s.segs.empty()) return OXR_OK;
    s.t0 = std::chrono::steady_clock::now();
    s.N = cursor;
    // the engine's config and tuning are read under its lock, so a concurrent
    // oxr_engine_set_config/set_tuning takes effect between steps, never inside one
    std::lock_guard<std::mutex> elk(e->mu);
    int rc = ensure_event_capacity(*e, s.buf, s.N);
    if (rc) return rc;
    int nb = oxl::compact_blocks(s.N);
    if (s.h_offs_n < size_t(nb) + 1) {
        if (s.h_offs) (void)hipHostFree(s.h_offs);
        s.h_offs = nullptr; s.h_offs_n = 0;
        OXR_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_offs), (size_t(nb) + 1) * sizeof(long long), hipHostMallocDefault));
        s.h_offs_n = size_t(nb) + 1;
    }
    OXR_HIP(s.d_segs.reserve(s.segs.size()));
    OXR_HIP(hipMemcpyAsync(s.d_segs.p, s.segs.data(), s.segs.size() * sizeof(oxd::Seg), hipMemcpyHostToDevice, s.stream));
    int form = cfg.form;
    int block = form == OXR_FORM_PERSISTENT ? e->tune.block_size : 256;
    oxd::Params P = e->params(s.buf, s.N, 0, 0, block);
    // served requests return sphoton hits only: no per-photon debug or photonlite arrays,
    // so the production kernels run whatever the engine's config asks of single launches
    P.seq = nullptr; P.record = nullptr; P.rec = nullptr; P.prd = nullptr; P.lite = nullptr;
    P.segs = s.d_segs.p; P.nseg = int(s.segs.size());
    OXR_HIP(hipMemsetAsync(s.buf.counters.p, 0, 8 * sizeof(unsigned long long), s.stream));
    OXR_HIP(hipEventRecord(s.e0, s.stream));
    if (oxl::needs_generate(P, false)) OXR_HIP(oxl::generate(s.stream, P));
    if (form == OXR_FORM_PERSISTENT) {
        int grid = std::max(1, e->tune.blocks_per_cu) * std::max(1, e->num_cu);
        OXR_HIP(oxl::simulate_persist(s.stream, P, block, grid, e->cfg.fp64_roots != 0, e->has_list));
    } else {
        OXR_HIP(oxl::simulate_plain(s.stream, P, block, e->cfg.fp64_roots != 0, false, e->has_list));
    }
    OXR_HIP(hipEventRecord(s.e1, s.stream));
    OXR_HIP(oxl::compact(s.stream, s.buf.hitflag.p, s.buf.photon.p, s.N, s.buf.counts.p, s.buf.offs.p, s.buf.hit.p, s.N));
    OXR_HIP(hipMemcpyAsync(s.h_offs, s.buf.offs.p, (size_t(nb) + 1) * sizeof(long long), hipMemcpyDeviceToHost, s.stream));
    for (size_t k = 0; k < s.segs.size(); k++) {
        s.owner[k]->dispatched += s.take[k];
        s.owner[k]->inflight += 1;
    }
    s.busy = true;
    s.seqno = next_seq++;
    return OXR_OK;
}

int Server::finalize_step(Slot& s) {
    const long long CH = 4096;
    s.busy = false;
    for (auto* r : s.owner) r->inflight -= 1;
    OXR_HIP(hipStreamSynchronize(s.stream));
    int nb = oxl::compact_blocks(s.N);
    long long nh = s.h_offs[nb];
    unsigned long long ctr[8] = {0};
    OXR_HIP(hipMemcpyAsync(ctr, s.buf.counters.p, sizeof(ctr), hipMemcpyDeviceToHost, s.stream));
    // each segment's hits: device-to-device into a block of its request; the slot's hit
    // buffer is reused by this slot's next step only after these copies (same stream)
    for (size_t k = 0; k < s.segs.size(); k++) {
        long long h0 = s.h_offs[size_t(s.segs[k].start / CH)];
        long long h1 = k + 1 < s.segs.size() ? s.h_offs[size_t(s.segs[k + 1].start / CH)] : nh;
        if (h1 <= h0) continue;
        HitBlock b{nullptr, 0, h1 - h0};
        OXR_HIP(pool.get(size_t(b.n) * 64, &b.p, &b.cap));
        s.owner[k]->hits.push_back(b);
        s.owner[k]->nhit += b.n;
        OXR_HIP(hipMemcpyAsync(b.p, s.buf.hit.p + 4 * h0, size_t(b.n) * 64, hipMemcpyDeviceToDevice, s.stream));
    }
    OXR_HIP(hipStreamSynchronize(s.stream));
    if (ctr[4] != 0) return fail(OXR_EINVAL, "traversal stack overflow");
    float kms = 0.f;
    (void)hipEventElapsedTime(&kms, s.e0, s.e1);
    auto w1 = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> lk(mu);
    stat_steps += 1;
    for (auto t : s.take) stat_photons += double(t);
    stat_busy_ms += std::chrono::duration<double, std::milli>(w1 - s.t0).count();
    stat_kernel_ms += kms;
    stat_bounce += double(ctr[0]);
    stat_hits += double(nh);
    return OXR_OK;
}

void Server::loop() {
    for (;;) {
        std::vector<std::unique_ptr<Req>> admitted;
        {
            std::unique_lock<std::mutex> lk(mu);
            bool inflight = slot[0].busy || slot[1].busy;
            cv_work.wait(lk, [&] { return stop || !incoming.empty() || !active.empty() || inflight; });
            if (stop) break;
            while (!incoming.empty()) {
                Req& r = *incoming.front();
                bool room = int(active.size() + admitted.size()) < cfg.max_active;
                bool mem = used + r.bytes <= budget || (active.empty() && admitted.empty());
                if (!room || !mem) break;
                used += r.bytes;
                r.t_admit = clock();
                admitted.push_back(std::move(incoming.front()));
                incoming.pop_front();
            }
        }
        for (auto& r : admitted) {
            int rc = admit_upload(*r);
            if (rc != OXR_OK || r->num_photon == 0) complete(r, rc);
            else active.push_back(std::move(r));
        }
        int rc = OXR_OK;
        Slot* fr = !slot[0].busy ? &slot[0] : (!slot[1].busy ? &slot[1] : nullptr);
        if (fr && dispatchable()) {
            rc = launch_step(*fr);
        } else if (slot[0].busy || slot[1].busy) {
            Slot* old = !slot[0].busy ? &slot[1] : (!slot[1].busy ? &slot[0] : (slot[0].seqno < slot[1].seqno ? &slot[0] : &slot[1]));
            rc = finalize_step(*old);
        } else {
            continue;
        }
        if (rc != OXR_OK) {
            for (auto& sl : slot) { if (sl.busy) { (void)hipStreamSynchronize(sl.stream); sl.busy = false; } }
            for (auto& r : active) complete(r, rc);
            active.clear();
            continue;
        }
        for (auto& r : active)
            if (r->dispatched >= r->num_photon && r->inflight == 0) complete(r, OXR_OK);
        active.erase(std::remove_if(active.begin(), active.end(), [](const std::unique_ptr<Req>& r) { return !r; }), active.end());
    }
    for (auto& sl : slot) if (sl.busy) { (void)hipStreamSynchronize(sl.stream); sl.busy = false; }
    for (auto& r : active) complete(r, OXR_ESTATE);
    active.clear();
}

} // namespace oxh

using namespace oxh;

extern "C" {

int oxr_server_create(oxr_engine* eh, const oxr_server_config* cfg, oxr_server** out) {
    Engine* e = reinterpret_cast<Engine*>(eh);
    if (!e || !out) return fail(OXR_EINVAL, "null argument");
    *out = nullptr;
    Server* s = new Server();
    s->e = e;
    if (cfg) s->cfg = *cfg;
    if (s->cfg.max_step_photons <= 0) s->cfg.max_step_photons = 4 << 20;
    if (s->cfg.hit_fraction <= 0.f) s->cfg.hit_fraction = 0.25f;
    if (s->cfg.form != OXR_FORM_PLAIN) s->cfg.form = OXR_FORM_PERSISTENT;
    if (s->cfg.max_active <= 0) s->cfg.max_active = 256;
    auto bail = [&](int rc) {
        for (auto& sl : s->slot) {
            sl.buf.release_all(); sl.d_segs.release();
            if (sl.h_offs) (void)hipHostFree(sl.h_offs);
            if (sl.e0) (void)hipEventDestroy(sl.e0);
            if (sl.e1) (void)hipEventDestroy(sl.e1);
            if (sl.stream) (void)hipStreamDestroy(sl.stream);
        }
        if (s->admit_stream) (void)hipStreamDestroy(s->admit_stream);
        delete s;
        return rc;
    };
    hipError_t he;
    if ((he = hipStreamCreateWithFlags(&s->admit_stream, hipStreamNonBlocking)) != hipSuccess) return bail(hip_fail(he, "server stream"));
    for (auto& sl : s->slot) {
        if ((he = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking)) != hipSuccess) return bail(hip_fail(he, "server stream"));
        if ((he = hipEventCreate(&sl.e0)) != hipSuccess || (he = hipEventCreate(&sl.e1)) != hipSuccess) return bail(hip_fail(he, "server events"));
        int rc;
        {
            std::lock_guard<std::mutex> elk(e->mu);
            rc = ensure_event_capacity(*e, sl.buf, s->cfg.max_step_photons + 4096 * 64);
        }
        if (rc) return bail(rc);
    }
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    s->budget = s->cfg.hbm_budget_bytes > 0 ? s->cfg.hbm_budget_bytes : int64_t(0.8 * double(fr));
    s->th = std::thread([s] { s->loop(); });
    *out = reinterpret_cast<oxr_server*>(s);
    return OXR_OK;
}

void oxr_server_destroy(oxr_server* sh) {
    Server* s = reinterpret_cast<Server*>(sh);
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    for (auto& r : s->incoming) s->release(*r);
    for (auto& kv : s->finished) { s->release(*kv.second); s->release_hits(*kv.second); }
    for (auto& sl : s->slot) {
        (void)hipStreamSynchronize(sl.stream);
        sl.buf.release_all(); sl.d_segs.release();
        if (sl.h_offs) (void)hipHostFree(sl.h_offs);
        (void)hipEventDestroy(sl.e0);
        (void)hipEventDestroy(sl.e1);
        (void)hipStreamDestroy(sl.stream);
    }
    (void)hipStreamDestroy(s->admit_stream);
    delete s;
}

double oxr_server_clock(oxr_server* sh) {
    Server* s = reinterpret_cast<Server*>(sh);
    return s ? s->clock() : 0.0;
}

int oxr_server_submit(oxr_server* sh, int64_t id, const float* gs, int64_t num_gs, int32_t event_index) {
    Server* s = reinterpret_cast<Server*>(sh);
    if (!s || num_gs < 0 || (num_gs > 0 && !gs)) return fail(OXR_EINVAL, "bad argument");
    for (int64_t g = 0; g < num_gs; g++) {
        uint32_t code;
        std::memcpy(&code, gs + 24 * g, 4);
        if (code == 19u) return fail(OXR_EINVAL, "INPUT_PHOTON gensteps are not served (no input photon array per request)");
    }
    auto r = std::make_unique<Req>();
    r->id = id;
    r->event_index = event_index;
    r->gs.assign(gs, gs + 24 * num_gs);
    r->num_gs = num_gs;
    r->t_submit = s->clock();
    int64_t n = 0;
    int rc = genstep_offsets(*s->e, r->gs.data(), num_gs, r->off, &n);
    if (rc) return rc;
    r->num_photon = n;
    r->bytes = 96 * num_gs + 8 * (num_gs + 1) + n * (4 + 64) + int64_t(double(s->cfg.hit_fraction) * 64.0 * double(n));
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->finished.count(id)) return fail(OXR_EINVAL, "request id already completed and not taken");
        s->incoming.push_back(std::move(r));
    }
    s->cv_work.notify_all();
    return OXR_OK;
}

int oxr_server_wait(oxr_server* sh, oxr_request_result* res, int32_t timeout_ms) {
    Server* s = reinterpret_cast<Server*>(sh);
    if (!s || !res) return fail(OXR_EINVAL, "bad argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] { return !s->done_ids.empty(); };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return fail(OXR_ETIMEOUT, "no request completed");
    int64_t id = s->done_ids.front();
    s->done_ids.pop_front();
    Req& r = *s->finished[id];
    res->request_id = r.id; res->event_index = r.event_index; res->status = r.status;
    res->num_photon = r.num_photon; res->num_hit = r.nhit; res->bounce_steps = 0;
    res->t_submit = r.t_submit; res->t_admit = r.t_admit; res->t_done = r.t_done;
    return OXR_OK;
}

int oxr_server_take_hits(oxr_server* sh, int64_t id, void* dst, int64_t cap, int64_t* count) {
    Server* s = reinterpret_cast<Server*>(sh);
    if (!s || !count) return fail(OXR_EINVAL, "bad argument");
    std::unique_ptr<Req> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->finished.find(id);
        if (it == s->finished.end()) return fail(OXR_ENOTFOUND, "unknown or not yet completed request");
        *count = it->second->nhit;
        if (cap < it->second->nhit) return fail(OXR_ECAPACITY, "destination smaller than hit count");
        if (it->second->nhit > 0 && !dst) return fail(OXR_EINVAL, "null destination");
        r = std::move(it->second);
        s->finished.erase(it);
    }
    // the blocks were filled on a serving stream that was synchronised before the request
    // completed; this copy runs on the caller's thread
    char* o = static_cast<char*>(dst);
    int rc = OXR_OK;
    for (auto& b : r->hits) {
        hipError_t e = hipMemcpy(o, b.p, size_t(b.n) * 64, hipMemcpyDeviceToHost);
        if (e != hipSuccess) { rc = hip_fail(e, "take_hits copy"); break; }
        o += size_t(b.n) * 64;
    }
    s->release_hits(*r);
    return rc;
}

int oxr_server_stats(oxr_server* sh, double* o) {
    Server* s = reinterpret_cast<Server*>(sh);
    if (!s || !o) return fail(OXR_EINVAL, "bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    o[0] = s->stat_steps; o[1] = s->stat_photons; o[2] = s->stat_busy_ms; o[3] = s->stat_kernel_ms;
    o[4] = s->stat_bounce; o[5] = s->stat_hits; o[6] = double(s->used); o[7] = double(s->budget);
    return OXR_OK;
}

} // extern "C"
