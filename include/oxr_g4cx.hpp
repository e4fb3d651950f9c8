/* oxr_g4cx.hpp -- C++ facade over the oxr C ABI with the calls eic-opticks' Geant4
 * applications make (src/GPUCerenkov.h, src/GPUPhotonSourceMinimal.h,
 * src/GPUPhotonFileSource.h):
 *
 *   G4CXOpticks::SetGeometry / Get / simulate / reset / simtrace / Finalize
 *       (g4cx/G4CXOpticks.hh:34-104); geometry comes from a CSGFoundry folder
 *       (the folder G4CXOpticks::SaveGeometry writes), not from a G4VPhysicalVolume;
 *   SEvt::AddGenstep / SetInputPhoton / GetNumHit / getHit / GetNumGenstepFromGenstep /
 *       GetNumPhotonCollected / getSimtrace (sysrap/SEvt.hh), one GPU event store;
 *   U4-style genstep makers taking the step values U4::CollectGenstep_* reads from
 *       G4Track/G4Step (u4/U4.cc:150-330) as plain numbers;
 *   load_photons_txt (src/GPUPhotonFileSource.h:47-88);
 *   Simulator<Base>: the launcher protocol of sysrap/SSimulator.h:16-37 (simulate_launch,
 *       simtrace_launch, launch, desc, simulate, simtrace, reset) over one engine.
 *
 * Header only; link with liboxr.so. Errors from the C ABI throw std::runtime_error with
 * oxr_last_error(). Environment read at SetGeometry, as SEventConfig reads it:
 * OPTICKS_MAX_BOUNCE, OPTICKS_MAX_SLOT (photons per launch; K/M/G/h/H/T prefixes as
 * sstr::ParseIntSpec), OPTICKS_HIT_MASK (integer mask).
 */
#pragma once

#include "oxr.h"

#include <cctype>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace oxr {

/* sysrap/sphoton.h:177-193, 64 bytes */
struct sphoton {
    float pos[3];
    float time;
    float mom[3];
    uint32_t hitcount_iindex;          /* hi16 hitcount, lo16 iindex */
    float pol[3];
    float wavelength;
    uint32_t orient_boundary_flag;     /* hi1 orient, 15 boundary, lo16 flag */
    uint32_t identity;
    uint32_t index;
    uint32_t flagmask;
    unsigned flag() const { return orient_boundary_flag & 0xffffu; }
    unsigned boundary() const { return (orient_boundary_flag & 0x7fff0000u) >> 16; }
    unsigned iindex() const { return hitcount_iindex & 0xffffu; }
    unsigned hitcount() const { return hitcount_iindex >> 16; }
};
static_assert(sizeof(sphoton) == 64, "sphoton is 64 bytes");

/* quad6 genstep, 96 bytes (sysrap/quad.h) */
struct quad6 {
    float f[24];
    uint32_t& u(int i) { return *reinterpret_cast<uint32_t*>(&f[i]); }
    int32_t& i(int k) { return *reinterpret_cast<int32_t*>(&f[k]); }
    uint32_t numphoton() const { return reinterpret_cast<const uint32_t*>(f)[3]; }
};
static_assert(sizeof(quad6) == 96, "quad6 is 96 bytes");

/* simtrace item, sevent::add_simtrace layout (sysrap/sevent.h:670-696) */
struct simtrace_item {
    float normal[3], t;
    float pos[3], tmin;
    float origin[3];
    uint32_t globalPrimIdx_boundary;
    float dir[3];
    uint32_t iindex_identity;
};
static_assert(sizeof(simtrace_item) == 64, "simtrace item is 64 bytes");

inline void check(int rc, const char* what) {
    if (rc != OXR_OK) throw std::runtime_error(std::string(what) + ": " + oxr_last_error());
}

/* sstr::ParseIntSpec (sysrap/sstr.h:1066-1113): optional scale prefix h K H M G T, X<n> = 2^n-1 */
inline int64_t ParseIntSpec(const char* spec) {
    if (!spec || !*spec) return 0;
    if (spec[0] == 'X') return int64_t((1ull << std::atol(spec + 1)) - 1ull);
    int64_t scale = 1;
    const char* e = spec;
    if (!std::isdigit(static_cast<unsigned char>(spec[0]))) {
        switch (spec[0]) {
            case 'h': scale = 100; break;
            case 'K': scale = 1000; break;
            case 'H': scale = 100000; break;
            case 'M': scale = 1000000; break;
            case 'G': scale = 1000000000; break;
            case 'T': scale = 1000000000000LL; break;
        }
        e = spec + 1;
    }
    return std::atoll(e) * scale;
}

/* One photon per line, "pos_x pos_y pos_z time mom_x mom_y mom_z pol_x pol_y pol_z
 * wavelength"; empty lines and lines starting '#' skipped, malformed lines skipped with a
 * warning on stderr (src/GPUPhotonFileSource.h:47-88). */
inline std::vector<sphoton> load_photons_txt(const std::string& path) {
    std::vector<sphoton> out;
    std::ifstream ifs(path);
    if (!ifs.is_open()) {
        std::cerr << "ERROR: cannot open photon file: " << path << std::endl;
        return out;
    }
    std::string line;
    int lineno = 0;
    while (std::getline(ifs, line)) {
        lineno++;
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        float v[11];
        bool ok = true;
        for (float& x : v) ok = ok && static_cast<bool>(ss >> x);
        if (!ok) {
            std::cerr << "WARNING: skipping malformed line " << lineno << ": " << line << std::endl;
            continue;
        }
        sphoton p{};
        p.pos[0] = v[0]; p.pos[1] = v[1]; p.pos[2] = v[2]; p.time = v[3];
        p.mom[0] = v[4]; p.mom[1] = v[5]; p.mom[2] = v[6];
        p.pol[0] = v[7]; p.pol[1] = v[8]; p.pol[2] = v[9]; p.wavelength = v[10];
        out.push_back(p);
    }
    return out;
}

/* ---- genstep makers (u4/U4.cc:150-330), step values passed in directly --------- */
struct CerenkovStep {            /* what MakeGenstep_G4Cerenkov_modified reads */
    int trackid = 0, matline = 0, numphoton = 0;
    double pos[3] = {0, 0, 0}, time = 0, delta_position[3] = {0, 0, 0}, step_length = 0;
    int code = 11;
    double charge = -1, weight = 1, pre_velocity = 0, post_velocity = 0;
    double beta_inverse = 1, wmin_nm = 60, wmax_nm = 820, max_cos = 0, max_sin2 = 0;
    double mean_photons1 = 0, mean_photons2 = 0;
};
inline quad6 MakeGenstep_G4Cerenkov_modified(const CerenkovStep& s) {
    quad6 g{};
    g.i(0) = 18;                       /* OpticksGenstep_G4Cerenkov_modified */
    g.i(1) = s.trackid; g.i(2) = s.matline; g.i(3) = s.numphoton;
    for (int k = 0; k < 3; k++) g.f[4 + k] = float(s.pos[k]);
    g.f[7] = float(s.time);
    for (int k = 0; k < 3; k++) g.f[8 + k] = float(s.delta_position[k]);
    g.f[11] = float(s.step_length);
    g.i(12) = s.code; g.f[13] = float(s.charge); g.f[14] = float(s.weight); g.f[15] = float(s.pre_velocity);
    g.f[16] = float(s.beta_inverse); g.f[17] = float(s.wmin_nm); g.f[18] = float(s.wmax_nm); g.f[19] = float(s.max_cos);
    g.f[20] = float(s.max_sin2); g.f[21] = float(s.mean_photons1); g.f[22] = float(s.mean_photons2);
    g.f[23] = float(s.post_velocity);
    return g;
}
struct ScintillationStep {       /* what MakeGenstep_DsG4Scintillation_r4695 reads */
    int trackid = 0, matline = 0, numphoton = 0;
    double pos[3] = {0, 0, 0}, time = 0, delta_position[3] = {0, 0, 0}, step_length = 0;
    int code = 11;
    double charge = -1, weight = 1, mean_velocity = 0;
    int scnt = 1;
    double scintillation_time = 0;
};
inline quad6 MakeGenstep_DsG4Scintillation_r4695(const ScintillationStep& s) {
    quad6 g{};
    g.i(0) = 5;                        /* OpticksGenstep_DsG4Scintillation_r4695 */
    g.i(1) = s.trackid; g.i(2) = s.matline; g.i(3) = s.numphoton;
    for (int k = 0; k < 3; k++) g.f[4 + k] = float(s.pos[k]);
    g.f[7] = float(s.time);
    for (int k = 0; k < 3; k++) g.f[8 + k] = float(s.delta_position[k]);
    g.f[11] = float(s.step_length);
    g.i(12) = s.code; g.f[13] = float(s.charge); g.f[14] = float(s.weight); g.f[15] = float(s.mean_velocity);
    g.i(16) = s.scnt;
    g.f[20] = float(s.scintillation_time);
    return g;
}

/* ---- SEvt: the GPU event's gensteps in, hits out --------------------------------- */
class SEvt {
public:
    static SEvt* Get_EGPU() { static SEvt e; return &e; }
    static SEvt* CreateOrReuse_EGPU() { return Get_EGPU(); }

    static int AddGenstep(const quad6& gs) { return Get_EGPU()->addGenstep(gs); }
    int addGenstep(const quad6& gs) {
        gensteps_.push_back(gs);
        return int(gensteps_.size()) - 1;
    }
    void SetInputPhoton(const std::vector<sphoton>& p) {
        input_ = p;
        quad6 g{};
        g.i(0) = 19; g.i(3) = int(p.size());      /* OpticksGenstep_INPUT_PHOTON */
        gensteps_.push_back(g);
    }
    int64_t GetNumGenstepFromGenstep(int) const { return int64_t(gensteps_.size()); }
    int64_t GetNumPhotonCollected(int) const {
        int64_t n = 0;
        for (const quad6& g : gensteps_) n += int64_t(reinterpret_cast<const uint32_t*>(g.f)[3]);
        return n;
    }
    unsigned GetNumHit(int) const { return unsigned(hits_.size()); }
    void getHit(sphoton& p, unsigned idx) const {
        if (idx >= hits_.size()) throw std::out_of_range("hit index");
        p = hits_[idx];
    }
    const std::vector<sphoton>& hits() const { return hits_; }
    const std::vector<simtrace_item>& simtrace() const { return simtrace_; }
    void clear() { gensteps_.clear(); input_.clear(); hits_.clear(); simtrace_.clear(); }

    std::vector<quad6> gensteps_;
    std::vector<sphoton> input_;
    std::vector<sphoton> hits_;
    std::vector<simtrace_item> simtrace_;
};

/* ---- G4CXOpticks: geometry on the GPU, simulate the SEvt's gensteps ------------- */
class G4CXOpticks {
public:
    static G4CXOpticks* Get() { return instance(); }

    /* CSGFoundry folder (cfbase or cfbase/CSGFoundry) in the reference layout */
    static G4CXOpticks* SetGeometry(const char* cfbase) {
        G4CXOpticks* g = instance();
        g->setGeometry(cfbase);
        return g;
    }
    static void Finalize() { instance()->release(); }

    void simulate(int eventID, bool reset_) {
        if (!engine_) throw std::runtime_error("G4CXOpticks::simulate before SetGeometry");
        SEvt* ev = SEvt::Get_EGPU();
        const int64_t n = ev->GetNumPhotonCollected(0);
        check(oxr_event_set_input_photons(engine_, ev->input_.empty() ? nullptr : reinterpret_cast<const float*>(ev->input_.data()),
                                           int64_t(ev->input_.size())), "set input photons");
        std::vector<sphoton> hits(size_t(n > 0 ? n : 0));
        int64_t num_hit = 0;
        oxr_launch_stats st{};
        check(oxr_simulate(engine_, reinterpret_cast<const float*>(ev->gensteps_.data()),
                           int64_t(ev->gensteps_.size()), eventID, OXR_FORM_PERSISTENT, hits.data(),
                           int64_t(hits.size()), &num_hit, &st), "simulate");
        hits.resize(size_t(num_hit));
        ev->hits_ = std::move(hits);
        last_ = st;
        if (reset_) reset(eventID);
    }
    /* drop the event's gensteps and input photons; hits stay readable until the next
     * simulate, as the reference keeps them until SEvt::clear */
    void reset(int) {
        SEvt* ev = SEvt::Get_EGPU();
        ev->gensteps_.clear();
        ev->input_.clear();
    }
    void simtrace(int) {
        if (!engine_) throw std::runtime_error("G4CXOpticks::simtrace before SetGeometry");
        SEvt* ev = SEvt::Get_EGPU();
        const float* gs = reinterpret_cast<const float*>(ev->gensteps_.data());
        int64_t n = 0;
        int rc = oxr_simtrace(engine_, gs, int64_t(ev->gensteps_.size()), nullptr, 0, &n);
        if (rc != OXR_OK && rc != OXR_ECAPACITY) check(rc, "simtrace");
        ev->simtrace_.assign(size_t(n), simtrace_item{});
        check(oxr_simtrace(engine_, gs, int64_t(ev->gensteps_.size()), reinterpret_cast<float*>(ev->simtrace_.data()),
                           n, &n), "simtrace");
    }
    const oxr_launch_stats& last_launch() const { return last_; }
    oxr_engine* engine() { return engine_; }
    const oxr_config& config() const { return cfg_; }

private:
    G4CXOpticks() = default;
    ~G4CXOpticks() { release(); }
    static G4CXOpticks* instance() { static G4CXOpticks g; return &g; }

    void release() {
        if (engine_) oxr_engine_destroy(engine_);
        engine_ = nullptr;
        if (foundry_) oxr_foundry_free(foundry_);
        foundry_ = nullptr;
    }
    void setGeometry(const char* cfbase) {
        release();
        check(oxr_foundry_load(cfbase, &foundry_), "load CSGFoundry");
        oxr_geometry geo{};
        oxr_physics phys{};
        check(oxr_foundry_arrays(foundry_, &geo, &phys), "CSGFoundry arrays");
        oxr_config_defaults(&cfg_);
        if (const char* v = std::getenv("OPTICKS_MAX_BOUNCE")) cfg_.max_bounce = int32_t(ParseIntSpec(v));
        if (const char* v = std::getenv("OPTICKS_MAX_SLOT")) {
            int64_t s = ParseIntSpec(v);
            if (s > 0) cfg_.max_photon = s;
        }
        if (const char* v = std::getenv("OPTICKS_HIT_MASK")) cfg_.hitmask = uint32_t(std::strtoul(v, nullptr, 0));
        check(oxr_engine_create(&geo, &phys, &cfg_, &engine_), "engine create");
    }

    oxr_foundry* foundry_ = nullptr;
    oxr_engine* engine_ = nullptr;
    oxr_config cfg_{};
    oxr_launch_stats last_{};
};

/* ---- Simulator: the launcher protocol of sysrap/SSimulator.h:16-37 -----------------
 * Base is the interface the host is built against: the reference's SSimulator, whose
 * pure virtual members these then override (oxr::Simulator<SSimulator>), or NoBase to
 * use the class on its own. simulate and simtrace take the gensteps of
 * SEvt::Get_EGPU() (the role QEvt plays for CSGOptiX, qudarap/QSim.cc:488-504) and leave
 * hits or simtrace items there; simulate_launch relaunches the engine's current event.
 * Every call returns the launch's wall seconds, as the reference's launch timing does.
 * Rendering (the OptiX render pipeline) is not part of this engine: render and
 * render_launch throw std::logic_error. The engine is borrowed, not owned. */
struct NoBase {
    virtual ~NoBase() = default;
};

template <class Base = NoBase>
class Simulator : public Base {
public:
    enum Mode { SIMULATE, SIMTRACE };   /* what launch() runs, the role of SEventConfig RGMode */

    explicit Simulator(oxr_engine* e, int32_t form = OXR_FORM_PERSISTENT) : engine_(e), form_(form) {
        if (!e) throw std::invalid_argument("oxr::Simulator needs an engine");
        desc_ = std::string(oxr_version()) + " : simulate and simtrace on HIP, "
                + (form == OXR_FORM_PERSISTENT ? "persistent" : "plain") + " form";
    }

    /* low level: launches over the engine's current event */
    double render_launch() { throw std::logic_error("oxr::Simulator: rendering is not part of this engine"); }
    double simtrace_launch() { return simtrace(event_); }
    double simulate_launch() {
        oxr_launch_stats st{};
        check(oxr_simulate_launch(engine_, form_, &st), "simulate launch");
        last_ = st;
        return st.total_ms * 1e-3;
    }
    double launch() { return mode_ == SIMTRACE ? simtrace_launch() : simulate_launch(); }
    const char* desc() const { return desc_.c_str(); }

    /* high level: the SEvt's gensteps in, hits (or simtrace items) out. Events larger
     * than the engine's max_photon run as several launches over genstep slices, each with
     * its photon slot offset, hits gathered after each (QSim::simulate, QSim.cc:455-526);
     * returns the summed launch seconds. */
    double simulate(int eventID, bool reset_ = false) {
        SEvt* ev = SEvt::Get_EGPU();
        event_ = eventID;
        // input photons belong to this event only: an event without them clears the
        // engine's copy, so INPUT_PHOTON gensteps cannot read an earlier event's photons
        check(oxr_event_set_input_photons(engine_, ev->input_.empty() ? nullptr : reinterpret_cast<const float*>(ev->input_.data()),
                                           int64_t(ev->input_.size())), "set input photons");
        oxr_config cfg{};
        check(oxr_engine_get_config(engine_, &cfg), "get config");
        const float* gs = reinterpret_cast<const float*>(ev->gensteps_.data());
        const int64_t ngs = int64_t(ev->gensteps_.size());
        ev->hits_.clear();
        double seconds = 0.0;
        uint64_t slot = 0;
        int64_t g0 = 0;
        do {
            int64_t g1 = g0, np = 0;
            while (g1 < ngs) {
                const int64_t k = int64_t(ev->gensteps_[size_t(g1)].numphoton());
                if (k > cfg.max_photon) throw std::runtime_error("oxr: a single genstep exceeds max_photon");
                if (np + k > cfg.max_photon) break;
                np += k;
                g1++;
            }
            int64_t n = 0;
            check(oxr_event_set_gensteps(engine_, gs + 24 * g0, g1 - g0, eventID, slot, &n), "set gensteps");
            seconds += simulate_launch();
            int64_t nh = 0;
            int rc = oxr_event_get(engine_, OXR_GET_HIT, nullptr, 0, &nh);
            if (rc != OXR_OK && rc != OXR_ECAPACITY) check(rc, "hit count");
            const size_t at = ev->hits_.size();
            ev->hits_.resize(at + size_t(nh));
            if (nh > 0) check(oxr_event_get(engine_, OXR_GET_HIT, ev->hits_.data() + at, nh, &nh), "gather hits");
            slot += uint64_t(n);
            g0 = g1;
        } while (g0 < ngs);
        if (reset_) reset(eventID);
        return seconds;
    }
    double simtrace(int eventID) {
        SEvt* ev = SEvt::Get_EGPU();
        event_ = eventID;
        const float* gs = reinterpret_cast<const float*>(ev->gensteps_.data());
        const auto t0 = std::chrono::steady_clock::now();
        int64_t n = 0;
        int rc = oxr_simtrace(engine_, gs, int64_t(ev->gensteps_.size()), nullptr, 0, &n);
        if (rc != OXR_OK && rc != OXR_ECAPACITY) check(rc, "simtrace");
        ev->simtrace_.assign(size_t(n), simtrace_item{});
        check(oxr_simtrace(engine_, gs, int64_t(ev->gensteps_.size()), reinterpret_cast<float*>(ev->simtrace_.data()),
                           n, &n), "simtrace");
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    double render(const char* = nullptr) { return render_launch(); }
    void reset(int) {
        SEvt* ev = SEvt::Get_EGPU();
        ev->gensteps_.clear();
        ev->input_.clear();
    }

    void set_mode(Mode m) { mode_ = m; }
    const oxr_launch_stats& last_launch() const { return last_; }

private:
    oxr_engine* engine_;
    int32_t form_;
    int event_ = 0;
    Mode mode_ = SIMULATE;
    std::string desc_;
    oxr_launch_stats last_{};
};

} // namespace oxr
