"""ctypes mirror of include/oxr.h structs plus helpers to marshal numpy arrays."""
from __future__ import annotations

import ctypes as C

import numpy as np

OXR_OK, OXR_EINVAL, OXR_ENOMEM, OXR_EHIP, OXR_ECAPACITY, OXR_ESTATE, OXR_ETIMEOUT, OXR_ENOTFOUND = 0, -1, -2, -3, -4, -5, -6, -7
OXR_FORM_PLAIN = 0
OXR_FORM_PERSISTENT = 1
OXR_RNG_PRODUCTION = 0
OXR_RNG_ALIGNED = 1
# oxr_simulate_noalloc limits (include/oxr.h)
OXR_NOALLOC_MAX_PHOTON = 1 << 19
OXR_NOALLOC_MAX_GENSTEP = 1 << 14
OXR_NOALLOC_SCRATCH_BYTES = 12716288
OXR_AUTOTUNE_COLUMNS = 11
(OXR_GET_PHOTON, OXR_GET_HIT, OXR_GET_SEQ, OXR_GET_RECORD, OXR_GET_SEED, OXR_GET_GENSTEP,
 OXR_GET_PHOTONLITE, OXR_GET_HITLITE, OXR_GET_REC, OXR_GET_PRD) = range(10)

# photon flags (sysrap/OpticksPhoton.h:22-46)
FLAGS = dict(CERENKOV=1 << 0, SCINTILLATION=1 << 1, TORCH=1 << 2, BULK_ABSORB=1 << 3,
             BULK_REEMIT=1 << 4, BULK_SCATTER=1 << 5, SURFACE_DETECT=1 << 6,
             SURFACE_ABSORB=1 << 7, SURFACE_DREFLECT=1 << 8, SURFACE_SREFLECT=1 << 9,
             BOUNDARY_REFLECT=1 << 10, BOUNDARY_TRANSMIT=1 << 11, NAN_ABORT=1 << 12,
             EFFICIENCY_COLLECT=1 << 13, EFFICIENCY_CULL=1 << 14, MISS=1 << 15)
ABBREV = dict(CERENKOV='CK', SCINTILLATION='SI', TORCH='TO', BULK_ABSORB='AB', BULK_REEMIT='RE',
              BULK_SCATTER='SC', SURFACE_DETECT='SD', SURFACE_ABSORB='SA', SURFACE_DREFLECT='DR',
              SURFACE_SREFLECT='SR', BOUNDARY_REFLECT='BR', BOUNDARY_TRANSMIT='BT', NAN_ABORT='NA',
              EFFICIENCY_COLLECT='EC', EFFICIENCY_CULL='EX', MISS='MI')
NIBBLE_ABBREV = {i + 1: ABBREV[k] for i, k in enumerate(FLAGS)}


class Geometry(C.Structure):
    _fields_ = [("solid", C.c_void_p), ("num_solid", C.c_int64),
                ("prim", C.c_void_p), ("num_prim", C.c_int64),
                ("node", C.c_void_p), ("num_node", C.c_int64),
                ("plan", C.c_void_p), ("num_plan", C.c_int64),
                ("itra", C.c_void_p), ("num_itra", C.c_int64),
                ("inst", C.c_void_p), ("num_inst", C.c_int64)]


class Physics(C.Structure):
    _fields_ = [("bnd", C.c_void_p), ("num_bnd", C.c_int32), ("num_wl", C.c_int32),
                ("wl_min", C.c_float), ("wl_step", C.c_float),
                ("optical", C.c_void_p),
                ("icdf", C.c_void_p), ("icdf_rows", C.c_int32), ("icdf_width", C.c_int32),
                ("hd_factor", C.c_int32), ("reemission_scint_time", C.c_float)]


class Config(C.Structure):
    _fields_ = [("max_bounce", C.c_int32), ("tmin", C.c_float), ("tmin0", C.c_float),
                ("eps0_mask", C.c_uint32), ("tmax", C.c_float), ("max_time", C.c_float),
                ("hitmask", C.c_uint32), ("rng_seed", C.c_uint64), ("rng_offset", C.c_uint64),
                ("skipahead_event", C.c_uint64), ("rng_schedule", C.c_int32),
                ("fp64_roots", C.c_int32), ("seq_enabled", C.c_int32), ("max_record", C.c_int32),
                ("max_photon", C.c_int64), ("photonlite", C.c_int32), ("max_rec", C.c_int32),
                ("max_prd", C.c_int32), ("extent_domain", C.c_float), ("time_domain", C.c_float),
                ("propagate_refine", C.c_uint32), ("refine_distance", C.c_float), ("vizmask", C.c_uint32),
                ("reserved", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("lds_top_nodes", C.c_int32), ("refill_batch", C.c_int32), ("stack_mode", C.c_int32),
                ("bvh_mode", C.c_int32), ("csg_stack", C.c_int32), ("bvh_radial", C.c_int32),
                ("photon_order", C.c_int32), ("generate_inplace", C.c_int32)]


class LaunchStats(C.Structure):
    _fields_ = [("num_photon", C.c_int64), ("num_hit", C.c_int64), ("bounce_steps", C.c_uint64),
                ("kernel_ms", C.c_double), ("compact_ms", C.c_double), ("total_ms", C.c_double),
                ("node_visits", C.c_uint64), ("prim_tests", C.c_uint64),
                ("csg_node_bytes", C.c_uint64), ("instance_visits", C.c_uint64), ("generate_ms", C.c_double)]


class EngineInfo(C.Structure):
    _fields_ = [("num_solid", C.c_int64), ("num_prim", C.c_int64), ("num_node", C.c_int64),
                ("num_inst", C.c_int64), ("tlas_nodes", C.c_int64), ("blas_nodes", C.c_int64),
                ("node_bytes", C.c_int32), ("num_cu", C.c_int32), ("device_bytes", C.c_int64),
                ("free_bytes", C.c_int64), ("total_bytes", C.c_int64),
                ("device_name", C.c_char * 64), ("arch", C.c_char * 32),
                ("csg_height", C.c_int32), ("csg_in_registers", C.c_int32), ("simple_leaves", C.c_int32),
                ("csg_in_lds", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("max_step_photons", C.c_int64), ("hbm_budget_bytes", C.c_int64),
                ("hit_fraction", C.c_float), ("form", C.c_int32), ("max_active", C.c_int32),
                ("headroom", C.c_int32)]


class RequestResult(C.Structure):
    _fields_ = [("request_id", C.c_int64), ("event_index", C.c_int32), ("status", C.c_int32),
                ("num_photon", C.c_int64), ("num_hit", C.c_int64), ("bounce_steps", C.c_uint64),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Marshal:
    """Keeps contiguous copies of the arrays alive while the C structs point at them."""

    def __init__(self, geometry, physics):
        self.keep = []

        def arr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            self.keep.append(a)
            return a

        g = geometry
        self.solid = arr(g['solid'], np.int32)
        self.prim = arr(g['prim'], np.float32)
        self.node = arr(g['node'], np.float32)
        self.plan = arr(g.get('plan', np.zeros((0, 4), np.float32)), np.float32)
        self.itra = arr(g.get('itra', np.zeros((0, 4, 4), np.float32)), np.float32)
        self.inst = arr(g['inst'], np.float32)
        self.geo = Geometry(ptr(self.solid), len(self.solid), ptr(self.prim), len(self.prim),
                            ptr(self.node), len(self.node), ptr(self.plan), len(self.plan),
                            ptr(self.itra), len(self.itra), ptr(self.inst), len(self.inst))
        p = physics
        self.bnd = arr(p['bnd'], np.float32)
        self.optical = arr(p['optical'], np.uint32)
        icdf = p.get('icdf')
        self.icdf = None if icdf is None else arr(np.asarray(icdf).reshape(np.asarray(icdf).shape[0], -1), np.float32)
        self.phys = Physics(ptr(self.bnd), self.bnd.shape[0], self.bnd.shape[3],
                            float(p.get('wl_min', 60.0)), float(p.get('wl_step', 1.0)),
                            ptr(self.optical), ptr(self.icdf),
                            0 if self.icdf is None else self.icdf.shape[0],
                            0 if self.icdf is None else self.icdf.shape[1],
                            int(p.get('hd_factor', 0)), float(p.get('reemission_scint_time', 0.0)))


def default_config(**kw):
    """Reference defaults (sysrap/SEventConfig.cc:45-106) with overrides."""
    c = Config()
    c.max_bounce = 31
    c.tmin = 0.05
    c.tmin0 = 0.05
    c.eps0_mask = FLAGS['TORCH'] | FLAGS['CERENKOV'] | FLAGS['SCINTILLATION'] | \
        FLAGS['BULK_SCATTER'] | FLAGS['BULK_REEMIT']
    c.tmax = 1e6
    c.max_time = 1e27
    c.hitmask = FLAGS['SURFACE_DETECT']
    c.rng_seed = 0
    c.rng_offset = 0
    c.skipahead_event = 100000
    c.rng_schedule = OXR_RNG_PRODUCTION
    c.fp64_roots = 0
    c.seq_enabled = 0
    c.max_record = 0
    c.max_photon = 0            # MaxSlot 0: sized from the device memory
    c.photonlite = 0
    c.max_rec = 0
    c.max_prd = 0
    c.extent_domain = 1000.0    # SEventConfig MaxExtentDomain (mm)
    c.time_domain = 10.0        # SEventConfig MaxTimeDomain (ns)
    c.propagate_refine = 0      # SEventConfig PropagateRefine
    c.refine_distance = 5000.0  # SEventConfig PropagateRefineDistance (mm)
    c.vizmask = 0xff            # Params vizmask (all eight instance visibility bits)
    c.reserved = 0
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def seqhis_str(seqhis_pair):
    """Decode a (2,) u64 seqhis into 'TO BT SD' form (sysrap/sseq.h:54-184)."""
    out = []
    for w in seqhis_pair:
        w = int(w)
        for s in range(16):
            nib = (w >> (4 * s)) & 0xf
            if nib == 0:
                return ' '.join(out)
            out.append(NIBBLE_ABBREV.get(nib, '?'))
    return ' '.join(out)
