"""Python binding of liboxr.so, the engine's C ABI (include/oxr.h).

The classes mirror the reference's host interface for the simulate path:

* ``Engine(geometry, physics, config)``  ~ ``CSGOptiX::Create(CSGFoundry*)`` + ``QSim`` upload
  (CSGOptiX/CSGOptiX.cc:367-399, qudarap/QSim.cc:118-250)
* ``Engine.set_gensteps / launch / get``  ~ ``QEvt::setGenstepUpload_NP`` /
  ``SSimulator::simulate_launch`` / ``QEvt::gather*`` (qudarap/QEvt.cc:332-445,934-1133)
* ``Engine.simulate(gs, event_index)``   ~ ``CSGOptiXService::simulate(gs, eventID)`` -> hits
  (CSGOptiX/CSGOptiXService.h:76-84)
* ``Server``                              continuous-batching serving loop

There is no fallback: without the compiled library every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get('OXR_LIB') or os.path.join(HERE, 'liboxr.so')

_lib = None


class OxrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f'oxr error {code}: {msg}')
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OxrError(-3, f'{LIB_PATH} is not built; run __graft_entry__.build()')
        L = C.CDLL(LIB_PATH)
        vp, i64, i32, u64 = C.c_void_p, C.c_int64, C.c_int32, C.c_uint64
        P = C.POINTER
        sig = {
            'oxr_version': (C.c_char_p, []),
            'oxr_last_error': (C.c_char_p, []),
            'oxr_config_defaults': (None, [P(abi.Config)]),
            'oxr_engine_create': (i32, [P(abi.Geometry), P(abi.Physics), P(abi.Config), P(vp)]),
            'oxr_engine_destroy': (None, [vp]),
            'oxr_engine_set_config': (i32, [vp, P(abi.Config)]),
            'oxr_engine_set_tuning': (i32, [vp, P(abi.Tuning)]),
            'oxr_engine_get_tuning': (i32, [vp, P(abi.Tuning)]),
            'oxr_engine_get_config': (i32, [vp, P(abi.Config)]),
            'oxr_engine_query': (i32, [vp, P(abi.EngineInfo)]),
            'oxr_event_set_gensteps': (i32, [vp, vp, i64, i32, u64, P(i64)]),
            'oxr_event_set_input_photons': (i32, [vp, vp, i64]),
            'oxr_event_set_precooked': (i32, [vp, vp, i64, i32]),
            'oxr_simulate_launch': (i32, [vp, i32, P(abi.LaunchStats)]),
            'oxr_simulate_launch_counting': (i32, [vp, P(abi.LaunchStats)]),
            'oxr_event_get': (i32, [vp, i32, vp, i64, P(i64)]),
            'oxr_event_get_merged': (i32, [vp, i32, C.c_float, vp, i64, P(i64)]),
            'oxr_merge_hits': (i32, [vp, i32, C.c_float, vp, i64, vp, i64, P(i64)]),
            'oxr_simulate': (i32, [vp, vp, i64, i32, i32, vp, i64, P(i64), P(abi.LaunchStats)]),
            'oxr_trace_rays': (i32, [vp, vp, i64, vp]),
            'oxr_simtrace': (i32, [vp, vp, i64, vp, i64, P(i64)]),
            'oxr_intersect_prim': (i32, [vp, i32, vp, i64, vp, vp]),
            'oxr_autotune': (i32, [vp, i32, vp, i32, P(i32)]),
            'oxr_server_create': (i32, [vp, P(abi.ServerConfig), P(vp)]),
            'oxr_server_destroy': (None, [vp]),
            'oxr_server_clock': (C.c_double, [vp]),
            'oxr_server_submit': (i32, [vp, i64, vp, i64, i32]),
            'oxr_server_wait': (i32, [vp, P(abi.RequestResult), i32]),
            'oxr_server_take_hits': (i32, [vp, i64, vp, i64, P(i64)]),
            'oxr_server_stats': (i32, [vp, vp]),
            'oxr_foundry_load': (i32, [C.c_char_p, P(vp)]),
            'oxr_foundry_arrays': (i32, [vp, P(abi.Geometry), P(abi.Physics)]),
            'oxr_foundry_free': (None, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise OxrError(rc, lib().oxr_last_error().decode())


def stats_dict(s):
    return {k: getattr(s, k) for k, _ in s._fields_}


class Engine:
    def __init__(self, geometry, physics, config=None, **overrides):
        self.m = abi.Marshal(geometry, physics)
        self.cfg = config if config is not None else abi.default_config(**overrides)
        h = C.c_void_p()
        _check(lib().oxr_engine_create(C.byref(self.m.geo), C.byref(self.m.phys), C.byref(self.cfg), C.byref(h)))
        self.h = h
        self.num_photon = 0

    @classmethod
    def from_scene(cls, scene, **overrides):
        kw = dict(scene.config)
        kw.update(overrides)
        return cls(scene.geometry, scene.physics, abi.default_config(**kw))

    def close(self):
        if getattr(self, 'h', None):
            lib().oxr_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- config -----------------------------------------------------------
    def config(self):
        """The engine's config as in force (max_photon 0 resolved from the device memory)."""
        c = abi.Config()
        _check(lib().oxr_engine_get_config(self.h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in c._fields_}

    def set_config(self, **kw):
        for k, v in kw.items():
            setattr(self.cfg, k, v)
        _check(lib().oxr_engine_set_config(self.h, C.byref(self.cfg)))

    def tuning(self):
        t = abi.Tuning()
        _check(lib().oxr_engine_get_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def set_tuning(self, **kw):
        t = abi.Tuning()
        _check(lib().oxr_engine_get_tuning(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        _check(lib().oxr_engine_set_tuning(self.h, C.byref(t)))

    def info(self):
        i = abi.EngineInfo()
        _check(lib().oxr_engine_query(self.h, C.byref(i)))
        d = {k: getattr(i, k) for k, _ in i._fields_}
        d['device_name'] = i.device_name.decode()
        d['arch'] = i.arch.decode()
        return d

    # -- event --------------------------------------------------------------
    def set_gensteps(self, gs, event_index=0, slot_offset=0):
        gs = np.ascontiguousarray(gs, dtype=np.float32)
        n = C.c_int64()
        _check(lib().oxr_event_set_gensteps(self.h, abi.ptr(gs), len(gs), event_index, slot_offset, C.byref(n)))
        self.num_photon = n.value
        return n.value

    def set_input_photons(self, photons):
        p = None if photons is None else np.ascontiguousarray(photons, dtype=np.float32)
        _check(lib().oxr_event_set_input_photons(self.h, abi.ptr(p), 0 if p is None else len(p)))

    def set_precooked(self, u):
        if u is None:
            _check(lib().oxr_event_set_precooked(self.h, None, 0, 0))
            return
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(len(u), -1)
        _check(lib().oxr_event_set_precooked(self.h, abi.ptr(u), u.shape[0], u.shape[1]))

    def launch(self, form=abi.OXR_FORM_PLAIN, counting=False):
        s = abi.LaunchStats()
        if counting:
            _check(lib().oxr_simulate_launch_counting(self.h, C.byref(s)))
        else:
            _check(lib().oxr_simulate_launch(self.h, form, C.byref(s)))
        return stats_dict(s)

    def get(self, what):
        n = C.c_int64()
        rc = lib().oxr_event_get(self.h, what, None, 0, C.byref(n))
        if rc not in (0, -4):
            _check(rc)
        shape, dt = {
            abi.OXR_GET_PHOTON: ((n.value, 4, 4), np.float32),
            abi.OXR_GET_HIT: ((n.value, 4, 4), np.float32),
            abi.OXR_GET_SEQ: ((n.value, 2, 2), np.uint64),
            abi.OXR_GET_RECORD: ((n.value, 4, 4), np.float32),
            abi.OXR_GET_SEED: ((n.value,), np.int32),
            abi.OXR_GET_GENSTEP: ((n.value, 6, 4), np.float32),
            abi.OXR_GET_PHOTONLITE: ((n.value, 4), np.uint32),
            abi.OXR_GET_HITLITE: ((n.value, 4), np.uint32),
            abi.OXR_GET_REC: ((n.value, 4), np.uint32),
            abi.OXR_GET_PRD: ((n.value, 2, 4), np.float32),
        }[what]
        out = np.zeros(shape, dtype=dt)
        m = C.c_int64()
        _check(lib().oxr_event_get(self.h, what, abi.ptr(out), n.value, C.byref(m)))
        if what == abi.OXR_GET_RECORD:
            out = out.reshape(self.num_photon, -1, 4, 4)
        elif what == abi.OXR_GET_REC:
            out = out.reshape(self.num_photon, -1, 4)
        elif what == abi.OXR_GET_PRD:
            out = out.reshape(self.num_photon, -1, 2, 4)
        return out

    def photons(self):
        return self.get(abi.OXR_GET_PHOTON)

    def merged(self, what=abi.OXR_GET_HIT, time_window=1.0):
        """Hits of the last launch merged by (identity, time bucket) on the device."""
        n = C.c_int64()
        rc = lib().oxr_event_get_merged(self.h, what, C.c_float(time_window), None, 0, C.byref(n))
        if rc not in (0, -4):
            _check(rc)
        shape = (n.value, 4, 4) if what == abi.OXR_GET_HIT else (n.value, 4)
        out = np.zeros(shape, dtype=np.float32 if what == abi.OXR_GET_HIT else np.uint32)
        m = C.c_int64()
        _check(lib().oxr_event_get_merged(self.h, what, C.c_float(time_window), abi.ptr(out), n.value, C.byref(m)))
        return out

    def merge_hits(self, items, what=abi.OXR_GET_HIT, time_window=1.0):
        """Merge an already selected host array (sphoton (n,4,4) or sphotonlite (n,4))."""
        a = np.ascontiguousarray(items)
        out = np.zeros_like(a)
        m = C.c_int64()
        _check(lib().oxr_merge_hits(self.h, what, C.c_float(time_window), abi.ptr(a), len(a), abi.ptr(out), len(a),
                                    C.byref(m)))
        return out[:m.value]

    def hits(self):
        return self.get(abi.OXR_GET_HIT)

    def seq(self):
        return self.get(abi.OXR_GET_SEQ)

    def seed(self):
        return self.get(abi.OXR_GET_SEED)

    def simulate(self, gs, event_index=0, form=abi.OXR_FORM_PLAIN, hit_capacity=None, out=None):
        """One request: gensteps in, hits out (sliced by config.max_photon). out, when
        given, is a reusable float32 (m, 4, 4) host array the hits are written into (the
        returned hits are a view of it); it needs room for every photon of the event."""
        gs = np.ascontiguousarray(gs, dtype=np.float32)
        from .gensteps import num_photons
        cap = num_photons(gs) if hit_capacity is None else hit_capacity
        if out is not None:
            hits = _host_out(out, cap)
        else:
            hits = np.zeros((max(cap, 1), 4, 4), dtype=np.float32)
        n = C.c_int64()
        s = abi.LaunchStats()
        _check(lib().oxr_simulate(self.h, abi.ptr(gs), len(gs), event_index, form, abi.ptr(hits), cap,
                                  C.byref(n), C.byref(s)))
        return hits[:n.value], stats_dict(s)

    # -- queries ------------------------------------------------------------
    def trace(self, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 2, 4)
        prd = np.zeros((len(rays), 2, 4), dtype=np.float32)
        _check(lib().oxr_trace_rays(self.h, abi.ptr(rays), len(rays), abi.ptr(prd)))
        return prd

    def simtrace(self, gensteps):
        """2-D cross-section rays of FRAME gensteps as the (n,4,4) simtrace array."""
        gs = np.ascontiguousarray(gensteps, dtype=np.float32).reshape(-1, 6, 4)
        n = C.c_int64(0)
        rc = lib().oxr_simtrace(self.h, abi.ptr(gs), len(gs), None, 0, C.byref(n))
        if rc not in (0, abi.OXR_ECAPACITY):
            _check(rc)
        out = np.zeros((n.value, 4, 4), dtype=np.float32)
        _check(lib().oxr_simtrace(self.h, abi.ptr(gs), len(gs), abi.ptr(out), n.value, C.byref(n)))
        return out

    def intersect_prim(self, prim, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 2, 4)
        isect = np.zeros((len(rays), 4), dtype=np.float32)
        valid = np.zeros(len(rays), dtype=np.int32)
        _check(lib().oxr_intersect_prim(self.h, prim, abi.ptr(rays), len(rays), abi.ptr(isect), abi.ptr(valid)))
        return isect, valid

    def autotune(self, repeat=2, max_results=256):
        res = np.zeros((max_results, abi.OXR_AUTOTUNE_COLUMNS), dtype=np.float64)
        n = C.c_int32()
        _check(lib().oxr_autotune(self.h, repeat, abi.ptr(res), max_results, C.byref(n)))
        return res[:min(n.value, max_results)], self.tuning()


def _host_out(out, cap):
    if not (isinstance(out, np.ndarray) and out.dtype == np.float32 and out.flags['C_CONTIGUOUS']
            and out.ndim == 3 and out.shape[1:] == (4, 4)):
        raise ValueError('out must be a C-contiguous float32 (m, 4, 4) array')
    if len(out) < max(cap, 1):
        raise ValueError(f'out holds {len(out)} rows, {cap} needed')
    return out


class Server:
    """Continuous-batching request loop over one resident engine."""

    def __init__(self, engine, max_step_photons=4 << 20, hbm_budget_bytes=0, hit_fraction=0.25,
                 form=abi.OXR_FORM_PERSISTENT, max_active=256, headroom=False):
        self.engine = engine
        cfg = abi.ServerConfig(max_step_photons, hbm_budget_bytes, hit_fraction, form, max_active, int(headroom))
        h = C.c_void_p()
        _check(lib().oxr_server_create(engine.h, C.byref(cfg), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, 'h', None):
            lib().oxr_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock(self):
        return lib().oxr_server_clock(self.h)

    def submit(self, request_id, gs, event_index):
        gs = np.ascontiguousarray(gs, dtype=np.float32)
        _check(lib().oxr_server_submit(self.h, request_id, abi.ptr(gs), len(gs), event_index))

    def wait(self, timeout_ms=-1):
        r = abi.RequestResult()
        rc = lib().oxr_server_wait(self.h, C.byref(r), timeout_ms)
        if rc == -6:
            return None
        _check(rc)
        return {k: getattr(r, k) for k, _ in r._fields_}

    def take_hits(self, request_id, num_hit, out=None):
        """The request's hits; into out (a reusable float32 (m, 4, 4) host array with
        m >= num_hit, returned as a view) when given."""
        out = _host_out(out, num_hit) if out is not None else np.empty((max(num_hit, 1), 4, 4), dtype=np.float32)
        n = C.c_int64()
        _check(lib().oxr_server_take_hits(self.h, request_id, abi.ptr(out), num_hit, C.byref(n)))
        return out[:n.value]

    def stats(self):
        o = np.zeros(8, dtype=np.float64)
        _check(lib().oxr_server_stats(self.h, abi.ptr(o)))
        return dict(steps=o[0], photons=o[1], busy_ms=o[2], kernel_ms=o[3], bounce_steps=o[4], hits=o[5],
                    used_bytes=o[6], budget_bytes=o[7])
