"""ctypes wrapper of the CPU oracle (liboxr_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker. The engine package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'liboxr_oracle.so')
sys.path.insert(0, os.path.join(os.path.dirname(HERE), 'eic-opticks_amd'))
from oxr import abi  # noqa: E402  (struct layouts only)

_lib = None


def build():
    subprocess.check_call(['make', '-s', '-C', HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, i64, i32, u64, u32 = C.c_void_p, C.c_int64, C.c_int32, C.c_uint64, C.c_uint32
        L.orc_philox4x32_10.argtypes = [vp, vp, vp]
        L.orc_uniforms.argtypes = [u64, u64, u64, u64, u64, i64, vp]
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(abi.Geometry), C.POINTER(abi.Physics), C.POINTER(abi.Config)]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_config.argtypes = [vp, C.POINTER(abi.Config)]
        L.orc_set_cull.argtypes = [vp, i32]
        L.orc_fill_seed.restype = i64
        L.orc_fill_seed.argtypes = [vp, i64, vp, i64]
        L.orc_intersect_prim.argtypes = [vp, i32, vp, i64, vp, vp, i32]
        L.orc_trace_rays.argtypes = [vp, vp, i64, vp, i32]
        L.orc_simtrace.argtypes = [vp, vp, i64, vp, i64]
        L.orc_xlog.argtypes = [vp, i64, vp]
        L.orc_simtrace.restype = i64
        L.orc_simulate.restype = i64
        L.orc_simulate.argtypes = [vp, vp, i64, i32, u64, vp, i32, vp, i32, i32, i64, i64,
                                   vp, vp, vp, i32, vp, vp, vp, vp, i32, vp, i32]
        L.orc_propagate_at_boundary.argtypes = [vp, C.c_float, C.c_float, vp, i32, i32, vp]
        L.orc_sample.argtypes = [i32, vp, i64, u64, vp]
        L.orc_boundary_lookup.argtypes = [vp, vp, i64, u32, u32, vp]
        L.orc_scint_wavelength.argtypes = [vp, vp, i64, vp]
        L.orc_lut.restype = i32
        L.orc_lut.argtypes = [i32, i32, i32, i32]
        _lib = L
    return _lib


P = abi.ptr


def philox(ctr, key):
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32)
    key = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(P(ctr), P(key), P(out))
    return out


def uniforms(n, seed=0, offset=0, skip=100000, event=0, photon=0):
    out = np.zeros(n, dtype=np.float32)
    lib().orc_uniforms(seed, offset, skip, event, photon, n, P(out))
    return out


def fill_seed(gs):
    gs = np.ascontiguousarray(gs, dtype=np.float32)
    n = lib().orc_fill_seed(P(gs), len(gs), None, 0)
    seed = np.zeros(n, dtype=np.int32)
    lib().orc_fill_seed(P(gs), len(gs), P(seed), n)
    return seed


def lut(op, a, b, acloser):
    return lib().orc_lut(op, a, b, int(acloser))


def propagate_at_boundary(mom, pol, normal, n1, n2, u):
    """Closed-form check entry: arrays (n,3); u (n,k) precooked uniforms."""
    mpn = np.ascontiguousarray(np.concatenate([mom, pol, normal], axis=1), dtype=np.float32)
    u = np.ascontiguousarray(u, dtype=np.float32)
    out = np.zeros((len(mpn), 7), dtype=np.float32)
    lib().orc_propagate_at_boundary(P(mpn), n1, n2, P(u), u.shape[1], len(mpn), P(out))
    return out[:, 0:3], out[:, 3:6], out[:, 6].astype(np.int64)


def sample(which, mom, pol, normal, n, seed=1):
    mpn = np.ascontiguousarray(np.concatenate([mom, pol, normal]), dtype=np.float32)
    out = np.zeros((n, 6), dtype=np.float32)
    lib().orc_sample({'rayleigh': 0, 'diffuse': 1, 'marsaglia': 2, 'specular': 3}[which],
                     P(mpn), n, seed, P(out))
    return out[:, 0:3], out[:, 3:6]


def time_bucket(time, tw):
    """unsigned(time / tw) in float32 with the saturating float->unsigned conversion."""
    q = (np.asarray(time, dtype=np.float32) / np.float32(tw)).astype(np.float32)
    out = np.zeros(q.shape, dtype=np.uint64)
    ok = q > 0
    big = q >= np.float32(4294967296.0)
    out[ok & ~big] = q[ok & ~big].astype(np.uint64)
    out[big] = 0xffffffff
    return out


def merge(items, time_window, mask=None):
    """SPM::merge_partial_select (sysrap/SPM.cu:135-260) restated with numpy loops.

    items: sphoton (n,4,4) f32 / (n,16) or sphotonlite (n,4) u32. mask: select items whose
    flagmask shares a bit with it (select_pred, sphoton.h:300-304, sphotonlite.h:93-97);
    None = already selected. Key (identity<<48 | unsigned(time/tw)) in 64-bit arithmetic
    (key_functor, sphoton.h:277-286, sphotonlite.h:68-77); stable sort by key; each run
    folded left to right with reduce_op (sphoton.h:288-298, sphotonlite.h:79-90)."""
    u = np.ascontiguousarray(items).view(np.uint32)
    u = u.reshape(len(u), -1).copy()
    lite = u.shape[1] == 4
    TIME, FM = (1, 3) if lite else (3, 15)
    if mask is not None:
        u = u[(u[:, FM] & np.uint32(mask)) != 0]
    if time_window == 0 or len(u) == 0:
        return u
    ident = (u[:, 0] & 0xffff) if lite else (u[:, 13] & 0x00ffffff)
    t = u[:, TIME].view(np.float32)
    key = (ident.astype(np.uint64) << np.uint64(48)) | time_bucket(t, time_window)
    order = np.argsort(key, kind='stable')
    out = []
    i = 0
    n = len(order)
    while i < n:
        j = i
        while j + 1 < n and key[order[j + 1]] == key[order[i]]:
            j += 1
        r = u[order[i]].copy()
        for k in range(i + 1, j + 1):
            b = u[order[k]]
            ra, rb = r[TIME:TIME + 1].view(np.float32)[0], b[TIME:TIME + 1].view(np.float32)[0]
            hc = ((int(r[0 if lite else 7]) >> 16) + (int(b[0 if lite else 7]) >> 16)) & 0xffff
            fm = int(r[FM]) | int(b[FM])
            if lite:
                r[TIME] = np.float32(np.fmin(ra, rb)).view(np.uint32)      # r = a; r.time = fminf
                r[0] = (hc << 16) | (int(r[0]) & 0xffff)                   # set_hitcount_identity
            else:
                if not (np.fmin(ra, rb) == ra):                            # r = a_first ? a : b
                    r = b.copy()
                r[7] = (int(r[7]) & 0xffff) | (hc << 16)                   # set_hitcount
            r[FM] = fm
        out.append(r)
        i = j + 1
    return np.array(out, dtype=np.uint32)


class Oracle:
    """Holds an oracle context over one scene's arrays."""

    def __init__(self, geometry, physics, config):
        self.m = abi.Marshal(geometry, physics)
        self.cfg = config
        self.h = lib().orc_create(C.byref(self.m.geo), C.byref(self.m.phys), C.byref(self.cfg))

    def __del__(self):
        try:
            if self.h:
                lib().orc_destroy(self.h)
        except Exception:
            pass

    def set_cull(self, on=True):
        lib().orc_set_cull(self.h, int(on))

    def set_config(self, cfg):
        self.cfg = cfg
        lib().orc_set_config(self.h, C.byref(cfg))

    def intersect_prim(self, prim, rays, precision=0):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 2, 4)
        isect = np.zeros((len(rays), 4), dtype=np.float32)
        valid = np.zeros(len(rays), dtype=np.int32)
        lib().orc_intersect_prim(self.h, prim, P(rays), len(rays), P(isect), P(valid), precision)
        return isect, valid

    def trace(self, rays, precision=0):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 2, 4)
        prd = np.zeros((len(rays), 2, 4), dtype=np.float32)
        lib().orc_trace_rays(self.h, P(rays), len(rays), P(prd), precision)
        return prd

    def simtrace(self, gs):
        gs = np.ascontiguousarray(gs, dtype=np.float32).reshape(-1, 6, 4)
        n = lib().orc_simtrace(self.h, P(gs), len(gs), None, 0)
        out = np.zeros((n, 4, 4), dtype=np.float32)
        lib().orc_simtrace(self.h, P(gs), len(gs), P(out), n)
        return out

    def simulate(self, gs, event_index=0, slot_offset=0, precooked=None, input_photons=None,
                 precision=0, nthreads=8, begin=0, end=-1, seq=True, max_record=0, lite=False,
                 max_rec=0, max_prd=0):
        gs = np.ascontiguousarray(gs, dtype=np.float32)
        total = len(fill_seed(gs))
        if end < 0 or end > total:
            end = total
        photon = np.zeros((total, 4, 4), dtype=np.float32)
        sq = np.zeros((total, 2, 2), dtype=np.uint64) if seq else None
        rec = np.zeros((total, max_record, 4, 4), dtype=np.float32) if max_record > 0 else None
        bounce = np.zeros(total, dtype=np.uint32)
        overflow = np.zeros(total, dtype=np.uint32)
        lt = np.zeros((total, 4), dtype=np.uint32) if lite else None
        rc = np.zeros((total, max_rec, 4), dtype=np.uint32) if max_rec > 0 else None
        pd = np.zeros((total, max_prd, 2, 4), dtype=np.float32) if max_prd > 0 else None
        pre = None
        npre = 0
        if precooked is not None:
            pre = np.ascontiguousarray(precooked, dtype=np.float32).reshape(len(precooked), -1)
            npre = pre.shape[1]
        inp = None if input_photons is None else np.ascontiguousarray(input_photons, dtype=np.float32)
        lib().orc_simulate(self.h, P(gs), len(gs), event_index, slot_offset, P(pre), npre, P(inp),
                           precision, nthreads, begin, end, P(photon), P(sq), P(rec), max_record,
                           P(bounce), P(overflow), P(lt), P(rc), max_rec, P(pd), max_prd)
        sl = slice(begin, end)
        return dict(photon=photon[sl], seq=None if sq is None else sq[sl],
                    record=None if rec is None else rec[sl], bounce=bounce[sl],
                    overflow=overflow[sl], total=total, lite=None if lt is None else lt[sl],
                    rec=None if rc is None else rc[sl], prd=None if pd is None else pd[sl])

    def boundary_lookup(self, nm, line, k):
        nm = np.ascontiguousarray(nm, dtype=np.float32)
        out = np.zeros((len(nm), 4), dtype=np.float32)
        lib().orc_boundary_lookup(self.h, P(nm), len(nm), line, k, P(out))
        return out

    def scint_wavelength(self, u):
        u = np.ascontiguousarray(u, dtype=np.float32)
        out = np.zeros(len(u), dtype=np.float32)
        lib().orc_scint_wavelength(self.h, P(u), len(u), P(out))
        return out


def xlog(x):
    """The photon loop's single-precision log, elementwise."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    lib().orc_xlog(P(x), x.size, P(out))
    return out
