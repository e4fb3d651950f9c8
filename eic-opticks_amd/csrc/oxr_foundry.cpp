// oxr_foundry.cpp -- reads a CSGFoundry folder as the reference persists it
// (CSGFoundry::save_, CSG/CSGFoundry.cc:2768-2800; CSGFoundry::load :2977-3050, loadArray :3324-3350; SSim under
// SSim/stree/standard, sysrap/SSim.hh:52-55) into host arrays that oxr_engine_create takes.
// Host only: no device calls, so it runs and is tested without a GPU.
//
//   <cfbase>/CSGFoundry/solid.npy  (ns,3,4) i4     prim.npy (np,4,4) f4   node.npy (nn,4,4) f4
//                       plan.npy   (nl,1,4) f4     itra.npy (nt,4,4) f4   inst.npy (ni,4,4) f4
//   <cfbase>/CSGFoundry/SSim/stree/standard/bnd.npy (nb,4,2,nw,4) f4 + bnd_meta.txt
//                       optical.npy (nb*4,4) u4/i4     icdf.npy (1|3,4096,1) f4 (optional)
//
// NPY files (format versions 1-3, little-endian, C order) are parsed here; the bnd
// wavelength domain comes from bnd_meta.txt "domain_low:" / "domain_step:" lines (NP
// metadata sidecar), defaulting to 60 nm / 1 nm (sysrap/sdomain.h).
#include "../../include/oxr.h"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace oxh {
int fail(int code, const std::string& msg);
}
using oxh::fail;

namespace {

struct Arr {
    std::vector<char> data;
    std::vector<int64_t> shape;
    char kind = 0;       // 'f', 'i', 'u'
    int itemsize = 0;
    bool present = false;
    int64_t dim0() const { return shape.empty() ? 0 : shape[0]; }
    int64_t count() const { int64_t n = 1; for (int64_t s : shape) n *= s; return n; }
};

bool exists(const std::string& p) { struct stat st; return stat(p.c_str(), &st) == 0; }

// parse "{'descr': '<f4', 'fortran_order': False, 'shape': (3, 4, 4), }"
int parse_header(const std::string& h, Arr& a, const std::string& path) {
    auto field = [&](const char* key) -> size_t {
        size_t k = h.find(key);
        if (k == std::string::npos) return k;
        k = h.find(':', k);
        return k == std::string::npos ? k : k + 1;
    };
    size_t d = field("'descr'");
    if (d == std::string::npos) return fail(OXR_EINVAL, path + ": NPY header without descr");
    size_t q0 = h.find('\'', d), q1 = q0 == std::string::npos ? q0 : h.find('\'', q0 + 1);
    if (q1 == std::string::npos) return fail(OXR_EINVAL, path + ": bad descr");
    std::string descr = h.substr(q0 + 1, q1 - q0 - 1);
    if (descr.size() < 3 || (descr[0] != '<' && descr[0] != '|'))
        return fail(OXR_EINVAL, path + ": only little-endian arrays are read (descr " + descr + ")");
    a.kind = descr[1];
    a.itemsize = std::atoi(descr.c_str() + 2);
    if (!((a.kind == 'f' || a.kind == 'i' || a.kind == 'u') && a.itemsize == 4))
        return fail(OXR_EINVAL, path + ": expected 4-byte float or integer items, got " + descr);
    size_t fo = field("'fortran_order'");
    if (fo != std::string::npos && h.compare(h.find_first_not_of(' ', fo), 4, "True") == 0)
        return fail(OXR_EINVAL, path + ": Fortran-ordered arrays are not read");
    size_t s = field("'shape'");
    if (s == std::string::npos) return fail(OXR_EINVAL, path + ": NPY header without shape");
    size_t p0 = h.find('(', s), p1 = p0 == std::string::npos ? p0 : h.find(')', p0);
    if (p1 == std::string::npos) return fail(OXR_EINVAL, path + ": bad shape");
    std::stringstream ss(h.substr(p0 + 1, p1 - p0 - 1));
    std::string tok;
    a.shape.clear();
    while (std::getline(ss, tok, ',')) {
        size_t b = tok.find_first_not_of(" \t");
        if (b == std::string::npos) continue;
        long long v = std::atoll(tok.c_str() + b);
        if (v < 0) return fail(OXR_EINVAL, path + ": negative dimension");
        a.shape.push_back(v);
    }
    return OXR_OK;
}

int read_npy(const std::string& path, Arr& a, bool required) {
    a = Arr();
    std::ifstream f(path, std::ios::binary);
    if (!f) return required ? fail(OXR_ENOTFOUND, path + ": not found") : OXR_OK;
    char magic[8];
    if (!f.read(magic, 8) || std::memcmp(magic, "\x93NUMPY", 6) != 0) return fail(OXR_EINVAL, path + ": not an NPY file");
    uint32_t hlen = 0;
    if (magic[6] == 1) {
        unsigned char b[2];
        if (!f.read(reinterpret_cast<char*>(b), 2)) return fail(OXR_EINVAL, path + ": truncated header");
        hlen = uint32_t(b[0]) | (uint32_t(b[1]) << 8);
    } else if (magic[6] == 2 || magic[6] == 3) {
        unsigned char b[4];
        if (!f.read(reinterpret_cast<char*>(b), 4)) return fail(OXR_EINVAL, path + ": truncated header");
        hlen = uint32_t(b[0]) | (uint32_t(b[1]) << 8) | (uint32_t(b[2]) << 16) | (uint32_t(b[3]) << 24);
    } else {
        return fail(OXR_EINVAL, path + ": unknown NPY version");
    }
    if (hlen > (1u << 20)) return fail(OXR_EINVAL, path + ": NPY header too long");
    std::string h(hlen, '\0');
    if (!f.read(&h[0], hlen)) return fail(OXR_EINVAL, path + ": truncated header");
    int rc = parse_header(h, a, path);
    if (rc) return rc;
    int64_t n = a.count();
    if (n < 0 || n > (int64_t(1) << 36)) return fail(OXR_EINVAL, path + ": unreasonable size");
    a.data.resize(size_t(n) * size_t(a.itemsize));
    if (n > 0 && !f.read(a.data.data(), std::streamsize(a.data.size()))) return fail(OXR_EINVAL, path + ": truncated data");
    a.present = true;
    return OXR_OK;
}

bool has_shape(const Arr& a, std::initializer_list<int64_t> tail) {
    if (a.shape.size() != tail.size() + 1) return false;
    size_t k = 1;
    for (int64_t t : tail) if (a.shape[k++] != t) return false;
    return true;
}

float meta_value(const std::string& path, const char* key, float fallback) {
    std::ifstream f(path);
    std::string line;
    const size_t kl = std::strlen(key);
    while (std::getline(f, line)) {
        if (line.compare(0, kl, key) == 0 && line.size() > kl && line[kl] == ':') return std::strtof(line.c_str() + kl + 1, nullptr);
    }
    return fallback;
}

} // namespace

struct oxr_foundry {
    Arr solid, prim, node, plan, itra, inst, bnd, optical, icdf;
    float wl_min = 60.f, wl_step = 1.f;
    int hd_factor = 0;
};

extern "C" {

int oxr_foundry_load(const char* cfbase, oxr_foundry** out) {
    if (!cfbase || !out) return fail(OXR_EINVAL, "null argument");
    *out = nullptr;
    std::string base(cfbase);
    std::string dir = exists(base + "/CSGFoundry/solid.npy") ? base + "/CSGFoundry" : base;
    if (!exists(dir + "/solid.npy")) return fail(OXR_ENOTFOUND, "no CSGFoundry folder (solid.npy) at " + base);
    auto* F = new oxr_foundry();
    auto bail = [&](int rc) { delete F; return rc; };
    int rc;
    if ((rc = read_npy(dir + "/solid.npy", F->solid, true)) || (rc = read_npy(dir + "/prim.npy", F->prim, true)) ||
        (rc = read_npy(dir + "/node.npy", F->node, true)) || (rc = read_npy(dir + "/plan.npy", F->plan, false)) ||
        (rc = read_npy(dir + "/itra.npy", F->itra, false)) || (rc = read_npy(dir + "/inst.npy", F->inst, true)))
        return bail(rc);
    std::string sdir = dir + "/SSim/stree/standard";
    if ((rc = read_npy(sdir + "/bnd.npy", F->bnd, true)) || (rc = read_npy(sdir + "/optical.npy", F->optical, true)) ||
        (rc = read_npy(sdir + "/icdf.npy", F->icdf, false)))
        return bail(rc);
    if (!has_shape(F->solid, {3, 4}) || F->solid.kind == 'f') return bail(fail(OXR_EINVAL, "solid.npy: expected (n,3,4) int32"));
    if (!has_shape(F->prim, {4, 4}) || F->prim.kind != 'f') return bail(fail(OXR_EINVAL, "prim.npy: expected (n,4,4) float32"));
    if (!has_shape(F->node, {4, 4}) || F->node.kind != 'f') return bail(fail(OXR_EINVAL, "node.npy: expected (n,4,4) float32"));
    if (!has_shape(F->inst, {4, 4}) || F->inst.kind != 'f') return bail(fail(OXR_EINVAL, "inst.npy: expected (n,4,4) float32"));
    if (F->plan.present && (!(has_shape(F->plan, {4}) || has_shape(F->plan, {1, 4})) || F->plan.kind != 'f'))
        return bail(fail(OXR_EINVAL, "plan.npy: expected (n,1,4) or (n,4) float32"));
    if (F->itra.present && (!has_shape(F->itra, {4, 4}) || F->itra.kind != 'f')) return bail(fail(OXR_EINVAL, "itra.npy: expected (n,4,4) float32"));
    if (F->bnd.shape.size() != 5 || F->bnd.shape[1] != 4 || F->bnd.shape[2] != 2 || F->bnd.shape[4] != 4 || F->bnd.kind != 'f')
        return bail(fail(OXR_EINVAL, "bnd.npy: expected (nb,4,2,nw,4) float32"));
    if (!has_shape(F->optical, {4}) || F->optical.kind == 'f' || F->optical.dim0() != 4 * F->bnd.dim0())
        return bail(fail(OXR_EINVAL, "optical.npy: expected (4*nb,4) 32-bit integers"));
    if (F->icdf.present) {
        bool ok = F->icdf.kind == 'f' && (F->icdf.shape.size() == 3 || F->icdf.shape.size() == 2) &&
                  (F->icdf.dim0() == 1 || F->icdf.dim0() == 3) && F->icdf.shape[1] >= 2 &&
                  (F->icdf.shape.size() == 2 || F->icdf.shape[2] == 1);
        if (!ok) return bail(fail(OXR_EINVAL, "icdf.npy: expected (1|3,width,1) float32"));
        F->hd_factor = F->icdf.dim0() == 3 ? 20 : 0;   // QSim::UploadComponents uses 20 with the 3-row ICDF
    }
    F->wl_min = meta_value(sdir + "/bnd_meta.txt", "domain_low", 60.f);
    F->wl_step = meta_value(sdir + "/bnd_meta.txt", "domain_step", 1.f);
    if (!(F->wl_step > 0.f)) return bail(fail(OXR_EINVAL, "bnd_meta.txt: domain_step must be positive"));
    *out = F;
    return OXR_OK;
}

int oxr_foundry_arrays(const oxr_foundry* F, oxr_geometry* g, oxr_physics* p) {
    if (!F || !g || !p) return fail(OXR_EINVAL, "null argument");
    auto fp = [](const Arr& a) { return a.present && !a.data.empty() ? reinterpret_cast<const float*>(a.data.data()) : nullptr; };
    g->solid = reinterpret_cast<const int32_t*>(F->solid.data.data()); g->num_solid = F->solid.dim0();
    g->prim = fp(F->prim); g->num_prim = F->prim.dim0();
    g->node = fp(F->node); g->num_node = F->node.dim0();
    g->plan = fp(F->plan); g->num_plan = F->plan.present ? F->plan.dim0() : 0;
    g->itra = fp(F->itra); g->num_itra = F->itra.present ? F->itra.dim0() : 0;
    g->inst = fp(F->inst); g->num_inst = F->inst.dim0();
    p->bnd = fp(F->bnd);
    p->num_bnd = int32_t(F->bnd.dim0());
    p->num_wl = int32_t(F->bnd.shape[3]);
    p->wl_min = F->wl_min;
    p->wl_step = F->wl_step;
    p->optical = reinterpret_cast<const uint32_t*>(F->optical.data.data());
    p->icdf = fp(F->icdf);
    p->icdf_rows = F->icdf.present ? int32_t(F->icdf.dim0()) : 0;
    p->icdf_width = F->icdf.present ? int32_t(F->icdf.shape[1]) : 0;
    p->hd_factor = F->hd_factor;
    p->reemission_scint_time = 0.f;
    return OXR_OK;
}

void oxr_foundry_free(oxr_foundry* F) { delete F; }

} // extern "C"
