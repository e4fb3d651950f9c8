// oxr_server.cpp -- serving form: a request loop that keeps geometry, tables and the RNG
// definition resident and batches genstep requests from many events into shared device
// steps (continuous batching).
//
//  * Admission: requests enter in arrival order while the active-slot count and the HBM
//    budget (64 B/photon state + 4 B seed + reserved hit bytes + gensteps) allow. A
//    request's gensteps are uploaded by the first step that holds its photons, from that
//    slot's pinned staging buffer on that slot's stream; a step on another slot that
//    holds the same request waits for the upload on the device.
//  * Steps: every step gives the active requests a share of the step's photon budget,
//    shortest remaining first; one persistent (or plain) launch covers all of them via
//    a segment table; segments start on compaction-chunk boundaries so each request's
//    hits are a contiguous range of the ordered hit array.
//  * Overlap: two step slots on two streams, so the next step is already queued when a
//    step drains and its workgroups fill the compute units the draining step leaves.
//    The loop never blocks on a stream: it polls each slot's completion event.
//  * Output without copy commands: a step's compaction writes its hits, chunk offsets
//    and launch counters straight into pinned host memory (a copy command queued behind
//    another slot's persistent kernel, which holds every wave slot, would wait for that
//    whole kernel; HIP also shares a few hardware queues among all streams, and two
//    streams on one in-order queue serialise). When the step's event completes its hits
//    are on the host, in a block the step's requests share; take_hits is a host memcpy.
//    The block is sized from the hit ratio seen so far; rows beyond it land in the
//    slot's device hit buffer and are fetched by a synchronous copy (rare).
//  * No allocation or free on the request path once the pools are warm: gensteps and
//    offsets live in blocks of a server-owned device pool (a hipFree would synchronise
//    the device and stall the overlapped steps), hits in blocks of a pinned host pool,
//    and served photons find their genstep by binary search of the offsets instead of
//    a per-photon seed array.
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

// Blocks in power-of-two size classes, kept for reuse until the server ends: device
// memory for request gensteps, pinned host memory for hits. A request takes the smallest
// free block of its class or above (device blocks at most four classes above); small
// device classes are carved from 64-block slabs, one hipMalloc each. A new pinned block
// costs milliseconds (pinning, then touching its pages once so the first host copy out
// of it takes no page faults), so any free pinned block that fits is used first. Used by
// the serving thread and by take_hits on the caller's thread.
template <bool HOST>
struct Pool {
    std::mutex m;
    std::multimap<size_t, void*> free_;
    std::vector<void*> all;
    ~Pool() { for (void* p : all) (void)(HOST ? hipHostFree(p) : hipFree(p)); }
    static size_t cls(size_t bytes) { size_t c = 256; while (c < bytes) c <<= 1; return c; }
    hipError_t get(size_t bytes, void** p, size_t* cap) {
        size_t c = cls(bytes);
        std::lock_guard<std::mutex> lk(m);
        auto it = free_.lower_bound(c);
        if (it != free_.end() && (HOST || it->first <= 4 * c)) {
            *p = it->second; *cap = it->first; free_.erase(it); return hipSuccess;
        }
        const int per = (!HOST && c <= (size_t(64) << 10)) ? 64 : 1;
        void* base = nullptr;
        hipError_t e = HOST ? hipHostMalloc(&base, c, hipHostMallocDefault) : hipMalloc(&base, c * size_t(per));
        if (e != hipSuccess) return e;
        all.push_back(base);
        if (HOST) std::memset(base, 0, c);
        for (int k = 1; k < per; k++) free_.emplace(c, static_cast<char*>(base) + size_t(k) * c);
        *p = base; *cap = c;
        return hipSuccess;
    }
    void put(void* p, size_t cap) {
        if (!p) return;
        std::lock_guard<std::mutex> lk(m);
        free_.emplace(cap, p);
    }
};

// One step's hits in pinned host memory, shared by the requests that had photons in it;
// back to the pool when the last of them has taken its hits.
struct HostHits {
    Pool<true>* pool = nullptr;
    void* p = nullptr;
    size_t cap = 0;
    ~HostHits() { if (pool) pool->put(p, cap); }
};
struct HitRef { std::shared_ptr<HostHits> blk; int64_t first, n; };

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
    int up_slot = -1;         // slot whose stream carries the upload of d_gs
    bool up_done = false;     // a step holding the request has finished (so has the upload)
    int64_t dispatched = 0;   // photons placed in launched steps
    int inflight = 0;         // launched steps holding its photons whose hits are not yet on the host
    std::vector<HitRef> hits; // in photon order
    int64_t nhit = 0;
    const float4* gs_dev() const { return static_cast<const float4*>(d_gs); }
    const long long* off_dev() const {
        return reinterpret_cast<const long long*>(static_cast<const char*>(d_gs) + size_t(num_gs) * 96);
    }
};

enum SlotState { SLOT_FREE, SLOT_RUNNING };

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, e_up = nullptr, e_done = nullptr;
    EventBuffers buf;
    int64_t cap = 0;                 // photons the slot's buffers hold
    DBuf<oxd::Seg> d_segs;
    oxd::Seg* h_segs = nullptr;      // pinned, max_active entries
    std::vector<Req*> owner;
    std::vector<int64_t> take;
    long long* h_offs = nullptr;     // pinned: chunk offsets (nb+1), then the 8 launch counters
    std::shared_ptr<HostHits> hits;  // pinned block the step's compaction writes
    int64_t hcap = 0;                // hit rows it holds
    char* stage = nullptr;           // pinned upload staging of the step's new requests
    size_t stage_cap = 0;
    int64_t N = 0;
    int nseg = 0;
    SlotState state = SLOT_FREE;
    uint64_t seqno = 0;
};

constexpr int NSLOT = 2;
constexpr long long CHUNK = 4096;    // compaction chunk (oxl::compact) a segment starts on

struct Server {
    Pool<false> pool;                // declared first: destroyed after every user below
    Pool<true> hpool;
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
    Slot slot[NSLOT];
    uint64_t next_seq = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now(), busy_t0;
    int nbusy = 0;
    double last_submit = -1.0;       // server clock of the latest submit (under mu)
    double prev_submit = -1.0;       // and of the one before it
    double hit_ratio = 0.25;         // hits per photon the next step's host block is sized for
    double stat_steps = 0, stat_photons = 0, stat_busy_ms = 0, stat_kernel_ms = 0, stat_bounce = 0, stat_hits = 0;
    double clock() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
    void loop();
    void release(Req& r) { pool.put(r.d_gs, r.d_cap); r.d_gs = nullptr; r.d_cap = 0; }
    int launch_step(int si);
    int advance(Slot& s, bool* moved);
    void complete(std::unique_ptr<Req>& r, int status);
    void fail_all(int rc);
    bool dispatchable() const {
        for (auto& r : active) if (r->dispatched < r->num_photon) return true;
        return false;
    }
    void mark_busy(int delta) {
        auto now = std::chrono::steady_clock::now();
        if (nbusy == 0 && delta > 0) busy_t0 = now;
        nbusy += delta;
        if (nbusy == 0 && delta < 0) {
            std::lock_guard<std::mutex> lk(mu);
            stat_busy_ms += std::chrono::duration<double, std::milli>(now - busy_t0).count();
        }
    }
};

void Server::complete(std::unique_ptr<Req>& r, int status) {
    r->status = status;
    r->t_done = clock();
    release(*r);
    if (status != OXR_OK) { r->hits.clear(); r->nhit = 0; }
    std::lock_guard<std::mutex> lk(mu);
    used -= r->bytes;
    done_ids.push_back(r->id);
    finished[r->id] = std::move(r);
    cv_done.notify_all();
}

int Server::launch_step(int si) {
    Slot& s = slot[si];
    std::vector<Req*> order;
    for (auto& r : active) if (r->dispatched < r->num_photon) order.push_back(r.get());
    std::sort(order.begin(), order.end(), [](const Req* a, const Req* b) {
        return (a->num_photon - a->dispatched) < (b->num_photon - b->dispatched);
    });
    s.owner.clear(); s.take.clear();
    long long cursor = 0;
    int64_t left = std::max<int64_t>(cfg.max_step_photons, CHUNK);
    int nseg = 0;
    for (Req* r : order) {
        if (left <= 0 || nseg >= cfg.max_active) break;
        int64_t t = std::min(left, r->num_photon - r->dispatched);
        t = std::min<int64_t>(t, s.cap - cursor);      // the slot's buffers bound the step
        if (t <= 0) break;
        oxd::Seg g{};
        g.start = cursor; g.count = t;
        g.photon0 = uint64_t(r->dispatched);
        g.num_gs = int(r->num_gs);      // gs and off once the request has its device block
        g.event_index = r->event_index;
        s.h_segs[nseg++] = g;
        s.owner.push_back(r); s.take.push_back(t);
        cursor += ((t + CHUNK - 1) / CHUNK) * CHUNK;
        left -= t;
    }
    if (nseg == 0) return OXR_OK;
    s.N = cursor;
    s.nseg = nseg;
    int nb = oxl::compact_blocks(s.N);
    // gensteps of requests no earlier step uploaded: staged here (the slot is free, so
    // its previous uploads have ended) and copied on this slot's stream
    size_t need = 0;
    for (Req* r : s.owner)
        if (!r->d_gs) need += (size_t(r->num_gs) * 96 + (size_t(r->num_gs) + 1) * 8 + 255) & ~size_t(255);
    if (need > s.stage_cap) {
        if (s.stage) OXR_HIP(hipHostFree(s.stage));
        s.stage = nullptr; s.stage_cap = 0;
        size_t want = std::max(need, size_t(4) << 20);
        OXR_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.stage), want, hipHostMallocDefault));
        s.stage_cap = want;
    }
    // the host block for the step's hits: the ratio seen so far plus a tenth
    {
        int64_t want = std::min<int64_t>(s.N, int64_t(double(s.N) * hit_ratio * 1.1) + CHUNK);
        auto blk = std::make_shared<HostHits>();
        OXR_HIP(hpool.get(size_t(want) * 64, &blk->p, &blk->cap));
        blk->pool = &hpool;
        s.hits = blk;
        s.hcap = int64_t(blk->cap / 64);
    }
    void* dhits = nullptr;
    OXR_HIP(hipHostGetDevicePointer(&dhits, s.hits->p, 0));
    void* doffs = nullptr;
    OXR_HIP(hipHostGetDevicePointer(&doffs, s.h_offs, 0));
    // the engine's config and tuning are read under its lock, so a concurrent
    // oxr_engine_set_config/set_tuning takes effect between steps, never inside one
    std::lock_guard<std::mutex> elk(e->mu);
    size_t at = 0;
    for (Req* r : s.owner) {
        if (!r->d_gs) {
            size_t gbytes = size_t(r->num_gs) * 96, obytes = (size_t(r->num_gs) + 1) * sizeof(long long);
            OXR_HIP(pool.get(gbytes + obytes, &r->d_gs, &r->d_cap));
            std::memcpy(s.stage + at, r->gs.data(), gbytes);
            std::memcpy(s.stage + at + gbytes, r->off.data(), obytes);
            OXR_HIP(hipMemcpyAsync(r->d_gs, s.stage + at, gbytes + obytes, hipMemcpyHostToDevice, s.stream));
            at += (gbytes + obytes + 255) & ~size_t(255);
            r->up_slot = si;
        } else if (!r->up_done && r->up_slot != si) {
            // uploaded by a step still in flight on another slot: its e_up (or a later
            // record of it, which comes after it on that stream) orders the upload first
            OXR_HIP(hipStreamWaitEvent(s.stream, slot[r->up_slot].e_up, 0));
        }
    }
    OXR_HIP(hipEventRecord(s.e_up, s.stream));
    for (int k = 0; k < nseg; k++) {
        s.h_segs[k].gs = s.owner[size_t(k)]->gs_dev();
        s.h_segs[k].off = s.owner[size_t(k)]->off_dev();
    }
    OXR_HIP(hipMemcpyAsync(s.d_segs.p, s.h_segs, size_t(nseg) * sizeof(oxd::Seg), hipMemcpyHostToDevice, s.stream));
    int form = cfg.form;
    int block = form == OXR_FORM_PERSISTENT ? e->tune.block_size : 256;
    oxd::Params P = e->params(s.buf, s.N, 0, 0, block);
    // served requests return sphoton hits only: no per-photon debug or photonlite arrays,
    // so the production kernels run whatever the engine's config asks of single launches
    P.seq = nullptr; P.record = nullptr; P.rec = nullptr; P.prd = nullptr; P.lite = nullptr;
    P.segs = s.d_segs.p; P.nseg = nseg;
    OXR_HIP(hipMemsetAsync(s.buf.counters.p, 0, 8 * sizeof(unsigned long long), s.stream));
    OXR_HIP(hipEventRecord(s.e0, s.stream));
    if (oxl::needs_generate(P, false)) OXR_HIP(oxl::generate(s.stream, P));
    if (form == OXR_FORM_PERSISTENT) {
        int per_cu = std::max(1, e->tune.blocks_per_cu);
        if (cfg.headroom) {
            int occ = oxl::persist_occupancy(P, block, e->cfg.fp64_roots != 0, e->has_list);
            if (occ > 0) per_cu = std::min(per_cu, occ);
            if (per_cu >= 2) per_cu -= 1;
        }
        int grid = per_cu * std::max(1, e->num_cu);
        OXR_HIP(oxl::simulate_persist(s.stream, P, block, grid, e->cfg.fp64_roots != 0, e->has_list));
    } else {
        OXR_HIP(oxl::simulate_plain(s.stream, P, block, e->cfg.fp64_roots != 0, false, e->has_list));
    }
    OXR_HIP(hipEventRecord(s.e1, s.stream));
    long long* hoffs = static_cast<long long*>(doffs);
    OXR_HIP(oxl::compact_host(s.stream, s.buf.hitflag.p, s.buf.photon.p, s.N, s.buf.counts.p, s.buf.offs.p, hoffs,
                              s.buf.counters.p, reinterpret_cast<unsigned long long*>(hoffs + nb + 1),
                              static_cast<float4*>(dhits), s.hcap, s.buf.hit.p, s.N));
    OXR_HIP(hipEventRecord(s.e_done, s.stream));
    for (int k = 0; k < nseg; k++) {
        s.owner[size_t(k)]->dispatched += s.take[size_t(k)];
        s.owner[size_t(k)]->inflight += 1;
    }
    s.state = SLOT_RUNNING;
    s.seqno = next_seq++;
    mark_busy(+1);
    return OXR_OK;
}

static bool event_done(hipEvent_t ev, hipError_t* err) {
    hipError_t q = hipEventQuery(ev);
    if (q == hipSuccess) return true;
    if (q != hipErrorNotReady) *err = q;
    return false;
}

// Retire a slot whose step has ended: its hits are already in the step's host block
// (rows beyond the block's capacity are fetched from the device buffer first); each
// segment's range goes to its request.
int Server::advance(Slot& s, bool* moved) {
    if (s.state != SLOT_RUNNING) return OXR_OK;
    hipError_t err = hipSuccess;
    if (!event_done(s.e_done, &err)) return err == hipSuccess ? OXR_OK : hip_fail(err, "server step");
    int nb = oxl::compact_blocks(s.N);
    const unsigned long long* ctr = reinterpret_cast<const unsigned long long*>(s.h_offs + nb + 1);
    if (ctr[4] != 0) return fail(OXR_EINVAL, "traversal stack overflow");
    long long nh = s.h_offs[nb];
    if (nh > s.hcap) {
        auto big = std::make_shared<HostHits>();
        OXR_HIP(hpool.get(size_t(nh) * 64, &big->p, &big->cap));
        big->pool = &hpool;
        std::memcpy(big->p, s.hits->p, size_t(s.hcap) * 64);
        OXR_HIP(hipMemcpyAsync(static_cast<char*>(big->p) + size_t(s.hcap) * 64, s.buf.hit.p + 4 * s.hcap,
                               size_t(nh - s.hcap) * 64, hipMemcpyDeviceToHost, s.stream));
        OXR_HIP(hipStreamSynchronize(s.stream));
        s.hits = big;
    }
    for (auto* r : s.owner) r->up_done = true;
    for (int k = 0; k < s.nseg; k++) {
        long long h0 = s.h_offs[size_t(s.h_segs[k].start / CHUNK)];
        long long h1 = k + 1 < s.nseg ? s.h_offs[size_t(s.h_segs[k + 1].start / CHUNK)] : nh;
        if (h1 <= h0) continue;
        s.owner[size_t(k)]->hits.push_back(HitRef{s.hits, h0, h1 - h0});
        s.owner[size_t(k)]->nhit += h1 - h0;
    }
    double ratio = s.N > 0 ? double(nh) / double(s.N) : 0.0;
    hit_ratio = std::max(ratio, 0.5 * (hit_ratio + ratio));
    float kms = 0.f;
    (void)hipEventElapsedTime(&kms, s.e0, s.e1);
    {
        std::lock_guard<std::mutex> lk(mu);
        stat_steps += 1;
        for (auto t : s.take) stat_photons += double(t);
        stat_kernel_ms += kms;
        stat_bounce += double(ctr[0]);
        stat_hits += double(nh);
    }
    for (auto* r : s.owner) r->inflight -= 1;
    s.owner.clear(); s.take.clear(); s.nseg = 0;
    s.hits.reset();
    s.state = SLOT_FREE;
    mark_busy(-1);
    *moved = true;
    return OXR_OK;
}

void Server::fail_all(int rc) {
    for (auto& sl : slot) {
        if (sl.state != SLOT_FREE) { (void)hipStreamSynchronize(sl.stream); mark_busy(-1); }
        sl.state = SLOT_FREE; sl.owner.clear(); sl.take.clear(); sl.nseg = 0; sl.hits.reset();
    }
    for (auto& r : active) { r->inflight = 0; complete(r, rc); }
    active.clear();
}

void Server::loop() {
    // While requests arrive in a burst (the latest two submits within COALESCE seconds),
    // a step that the pending photons would not fill waits up to COALESCE seconds after
    // the latest submit (2 x COALESCE at most in all), so the burst is served in full
    // steps rather than a small one and a full one. A lone request launches at once.
    constexpr double COALESCE = 5e-4;
    bool idle = false;
    double hold = 0.0;               // > 0: hold the launch until this server clock
    double hold_since = -1.0;        // when the current hold began
    for (;;) {
        std::vector<std::unique_ptr<Req>> admitted;
        double submitted, submitted_prev;
        {
            std::unique_lock<std::mutex> lk(mu);
            auto wake = [&] { return stop || !incoming.empty(); };
            if (hold > 0.0) {
                double left = hold - clock();
                if (left > 0) cv_work.wait_for(lk, std::chrono::duration<double>(left), wake);
            } else if (idle) {
                // device work in flight: poll its events every 50 us, waking at once for
                // a new request; nothing in flight: sleep until a request or stop
                if (nbusy > 0) cv_work.wait_for(lk, std::chrono::microseconds(50), wake);
                else cv_work.wait(lk, wake);
            }
            if (stop) break;
            submitted = last_submit;
            submitted_prev = prev_submit;
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
        bool moved = !admitted.empty();
        int rc = OXR_OK;
        for (auto& r : admitted) {
            if (r->num_photon == 0) complete(r, OXR_OK);
            else active.push_back(std::move(r));
        }
        // oldest step first, so hits reach each request in photon order
        Slot* order[NSLOT];
        for (int k = 0; k < NSLOT; k++) order[k] = &slot[k];
        std::sort(order, order + NSLOT, [](const Slot* a, const Slot* b) { return a->seqno < b->seqno; });
        for (int k = 0; k < NSLOT && rc == OXR_OK; k++) {
            if (order[k]->state == SLOT_FREE) continue;
            rc = advance(*order[k], &moved);
            if (order[k]->state == SLOT_RUNNING) break;   // later steps retire after it
        }
        if (rc == OXR_OK) {
            for (auto& r : active)
                if (r->dispatched >= r->num_photon && r->inflight == 0) complete(r, OXR_OK);
            active.erase(std::remove_if(active.begin(), active.end(), [](const std::unique_ptr<Req>& r) { return !r; }), active.end());
            hold = 0.0;
            if (dispatchable()) {
                int64_t pending = 0;
                for (auto& r : active) pending += r->num_photon - r->dispatched;
                double now = clock();
                if (hold_since < 0.0) hold_since = now;
                bool burst = submitted_prev >= 0.0 && submitted - submitted_prev < COALESCE;
                if (burst && pending < cfg.max_step_photons && now - submitted < COALESCE && now - hold_since < 2 * COALESCE)
                    hold = std::min(submitted + COALESCE, hold_since + 2 * COALESCE);
            }
            if (hold == 0.0) hold_since = -1.0;
            for (int si = 0; si < NSLOT && hold == 0.0; si++) {
                if (slot[si].state != SLOT_FREE || !dispatchable()) continue;
                rc = launch_step(si);
                if (rc != OXR_OK) break;
                moved = true;
            }
        }
        if (rc != OXR_OK) { fail_all(rc); moved = true; }
        idle = !moved;
    }
    for (auto& sl : slot) if (sl.state != SLOT_FREE) { (void)hipStreamSynchronize(sl.stream); sl.state = SLOT_FREE; }
    for (auto& r : active) { r->inflight = 0; complete(r, OXR_ESTATE); }
    active.clear();
}

} // namespace oxh

using namespace oxh;

extern "C" {

static void destroy_server(Server* s) {
    for (auto& sl : s->slot) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        sl.buf.release_all(); sl.d_segs.release();
        if (sl.h_offs) (void)hipHostFree(sl.h_offs);
        if (sl.h_segs) (void)hipHostFree(sl.h_segs);
        if (sl.stage) (void)hipHostFree(sl.stage);
        sl.hits.reset();
        for (hipEvent_t ev : {sl.e0, sl.e1, sl.e_up, sl.e_done}) if (ev) (void)hipEventDestroy(ev);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    delete s;
}

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
    auto bail = [&](int rc) { destroy_server(s); return rc; };
    hipError_t he;
    // a step holds at most max_step_photons plus one partial chunk per segment
    const int64_t cap = s->cfg.max_step_photons + int64_t(s->cfg.max_active) * CHUNK;
    for (auto& sl : s->slot) {
        if ((he = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking)) != hipSuccess) return bail(hip_fail(he, "server stream"));
        if ((he = hipEventCreate(&sl.e0)) != hipSuccess || (he = hipEventCreate(&sl.e1)) != hipSuccess ||
            (he = hipEventCreateWithFlags(&sl.e_up, hipEventDisableTiming)) != hipSuccess ||
            (he = hipEventCreateWithFlags(&sl.e_done, hipEventDisableTiming)) != hipSuccess)
            return bail(hip_fail(he, "server events"));
        int rc;
        {
            std::lock_guard<std::mutex> elk(e->mu);
            rc = ensure_event_capacity(*e, sl.buf, cap);
        }
        if (rc) return bail(rc);
        sl.cap = cap;
        int nb = oxl::compact_blocks(cap);
        if ((he = sl.d_segs.reserve(size_t(s->cfg.max_active))) != hipSuccess ||
            (he = hipHostMalloc(reinterpret_cast<void**>(&sl.h_segs), size_t(s->cfg.max_active) * sizeof(oxd::Seg), hipHostMallocDefault)) != hipSuccess ||
            (he = hipHostMalloc(reinterpret_cast<void**>(&sl.h_offs), (size_t(nb) + 1 + 8) * sizeof(long long), hipHostMallocDefault)) != hipSuccess)
            return bail(hip_fail(he, "server step buffers"));
    }
    // warm the pinned hit pool with blocks of the class a full step at the configured hit
    // fraction asks for: one per slot and three held by requests not yet taken
    s->hit_ratio = double(s->cfg.hit_fraction);
    {
        constexpr int NWARM = NSLOT + 3;
        size_t want = size_t(double(cap) * s->hit_ratio * 1.1 + double(CHUNK)) * 64;
        void* p[NWARM] = {nullptr};
        size_t c[NWARM] = {0};
        for (int k = 0; k < NWARM; k++)
            if ((he = s->hpool.get(want, &p[k], &c[k])) != hipSuccess) return bail(hip_fail(he, "server hit pool"));
        for (int k = 0; k < NWARM; k++) s->hpool.put(p[k], c[k]);
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
    for (auto& kv : s->finished) { s->release(*kv.second); kv.second->hits.clear(); }
    s->incoming.clear();
    s->finished.clear();
    destroy_server(s);
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
        s->prev_submit = s->last_submit;
        s->last_submit = s->clock();
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
    // the hits reached pinned host memory before the request completed: a host copy,
    // no device work on the caller's thread; large ones split over a few threads (one
    // thread copies at about a third of what the memory system delivers)
    struct Piece { const char* src; char* dst; size_t n; };
    std::vector<Piece> pieces;
    char* o = static_cast<char*>(dst);
    size_t total = 0;
    for (auto& h : r->hits) {
        pieces.push_back({static_cast<const char*>(h.blk->p) + size_t(h.first) * 64, o, size_t(h.n) * 64});
        o += size_t(h.n) * 64;
        total += size_t(h.n) * 64;
    }
    const size_t SPLIT = size_t(16) << 20;
    int nt = int(std::min<size_t>(8, total / SPLIT));
    if (nt <= 1) {
        for (auto& p : pieces) std::memcpy(p.dst, p.src, p.n);
    } else {
        // thread t copies bytes [t*total/nt, (t+1)*total/nt) of the concatenated pieces
        auto work = [&](int t) {
            size_t b = total * size_t(t) / size_t(nt), e = total * size_t(t + 1) / size_t(nt), at = 0;
            for (auto& p : pieces) {
                size_t lo = std::max(b, at), hi = std::min(e, at + p.n);
                if (lo < hi) std::memcpy(p.dst + (lo - at), p.src + (lo - at), hi - lo);
                at += p.n;
            }
        };
        std::vector<std::thread> th;
        for (int t = 1; t < nt; t++) th.emplace_back(work, t);
        work(0);
        for (auto& x : th) x.join();
    }
    r->hits.clear();
    return OXR_OK;
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
